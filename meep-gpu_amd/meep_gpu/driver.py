"""FdtdDriver: the reference driver's API over the HIP engine.

Mirrors reference ``meep_gpu.driver.FdtdDriver`` (driver.py:796) for the configurations the
engine steps: build the material (``set_epsilon_components`` :1220), poles
(``add_susceptibility`` :1531), the absorber (``setup_pml`` :2155), sources (``add_source``
:2567), flux monitors (``add_flux_monitor`` :3001), then ``step`` (:3419) / ``run`` (:3667).

The configuration freezes at the first step, as the reference's does (driver.py:3505-3521):
the device simulation is created then, with every table uploaded once (source currents for
the whole planned run included), and from there a step is two fused launches and the flux
sampling, with no host work in between. ``fast_path_report`` names the engine that ran.
"""

from __future__ import annotations

import math
from dataclasses import dataclass
from typing import Dict, List, Optional, Sequence, Tuple

import numpy

from .absorber import AbsorbingFace, graded_sigma, faces_for
from .chunks import IYEE, ChunkDivision
from .dft import DftFieldsMonitor, FluxMonitor, meep_frequency_span
from .dispersion import Susceptibility
from .engine import HipSimulation
from .fastpath import FastPathError, last_dispatch_report, plan_fast_path
from .grid import Grid
from .pml import PML
from .readback import meep_array
from .sources import (DepositBox, GaussianSource, current_table, meep_volume_source,
                      polarization_boxes)
from .workloads import Workload

_E = ("Ex", "Ey", "Ez")


def _boundary_kinds(spec) -> Dict[int, str]:
    """Per-axis boundaries from the reference's spellings (grid.py ``_normalize_boundaries``):
    None, one kind for every axis, an ``{"x": kind}`` mapping, or a 3-sequence, each entry a
    kind or a ``(low, high)`` pair whose sides agree."""
    if spec is None or (not isinstance(spec, str) and len(spec) == 0):
        return {}
    if isinstance(spec, str):
        items = {a: spec for a in range(3)}
    elif isinstance(spec, dict):
        items = {(k if isinstance(k, int) else "xyz".index(str(k).lower())): v
                 for k, v in spec.items()}
    else:
        if len(spec) != 3:
            raise ValueError(f"boundaries as a sequence take one entry per axis, got {spec!r}")
        items = dict(enumerate(spec))
    kinds = {}
    for axis, kind in items.items():
        if not isinstance(kind, str):
            low, high = kind
            if low != high:
                raise ValueError(f"axis {'xyz'[axis]}: the two faces must agree, got {kind!r}")
            kind = low
        kinds[axis] = kind
    return kinds


def _mirror_entries(symmetry) -> Tuple[Tuple[str, int], ...]:
    """``(axis, phase)`` per mirror plane: a bare axis name (an even plane), an ``(axis,
    phase)`` pair, or an object with ``axis`` (or ``direction``) and ``phase`` attributes, as
    ``mp.Mirror(direction, phase)`` and the reference's ``Mirror`` carry them."""
    out = []
    for entry in symmetry or ():
        if isinstance(entry, (str, int)):
            out.append((entry, 1))
        elif isinstance(entry, (tuple, list)):
            out.append((entry[0], int(entry[1])))
        else:
            axis = getattr(entry, "axis", getattr(entry, "direction", None))
            out.append((axis, int(getattr(entry, "phase", 1))))
    return tuple(out)


def bloch_phase(k_value: float, n_full: int, resolution: float) -> Optional[complex]:
    """The factor a field picks up one lattice vector up an axis, MEEP's ``eikna`` (reference
    grid.py:1009-1030, boundaries.cpp ``use_bloch``): None at k = 0, exactly -1 at the zone
    edge ``k * n == resolution / 2``, else ``exp(i * k * ((2*pi/a) * n))`` with the product
    associated as MEEP forms it."""
    if k_value == 0.0:
        return None
    if k_value * n_full == 0.5 * resolution:
        return complex(-1.0, 0.0)
    angle = k_value * ((2 * math.pi / resolution) * n_full)
    return complex(math.cos(angle), math.sin(angle))


class FdtdDriver:
    def __init__(self, cell_size: Sequence[float], resolution: float, courant: float = 0.5,
                 force_complex_fields: bool = False, symmetry: Sequence = (),
                 k_point: Sequence[float] = (0.0, 0.0, 0.0), boundaries=None,
                 dimensions: int = 3, prefer_gpu: bool = True, max_steps: int = 100000):
        """The reference's arguments (driver.py:811-827) in its order. Two defaults differ on
        purpose: real storage (``force_complex_fields=False``, MEEP's own default; the
        reference defaults to complex), and the GPU (``prefer_gpu``; there is no array path).
        A real run is the real part of the same run on complex storage, word for word.

        ``symmetry``: mirror planes (``("X", -1)``, a bare ``"Y"``, or ``mp.Mirror``-like
        objects); each halves its axis as MEEP does. ``k_point``: the Bloch wavevector in
        MEEP's units; a nonzero component needs complex fields and is refused on a mirrored,
        invariant, metallic or absorbing axis, as the reference refuses it
        (driver.py:1094-1170, grid.py:960-1000). ``boundaries``: per axis "periodic" (the
        default) or "metallic" (``set_boundaries``)."""
        if not prefer_gpu:
            raise ValueError("this package steps on the GPU only; prefer_gpu=False has no engine")
        self._cell = tuple(float(v) for v in cell_size)
        self._resolution, self._courant, self._dimensions = resolution, courant, dimensions
        self.force_complex_fields = bool(force_complex_fields)
        self.symmetry = _mirror_entries(symmetry)
        k = tuple(float(v) for v in k_point)
        if len(k) != 3 or not all(math.isfinite(v) for v in k):
            raise ValueError(f"k_point takes three finite real components, got {k_point!r}")
        self.k_point = k
        self._boundaries: Dict[int, str] = {}
        self.grid = self._make_grid()
        bloch = [a for a in range(3) if k[a] != 0.0]
        if bloch and not self.force_complex_fields:
            raise ValueError(f"k_point={k} is nonzero on {['xyz'[a] for a in bloch]}: the "
                             "boundary phase needs complex fields (force_complex_fields=True); "
                             "MEEP refuses the same pairing")
        for a in bloch:
            if self.grid.is_invariant(a):
                raise ValueError(f"k_point is nonzero on the invariant axis {'xyz'[a]}")
            if self.grid.folds[a] is not None:
                raise ValueError(f"k_point={k} is nonzero on the mirror-symmetric axis "
                                 f"{'xyz'[a]}: inside the zone a mirror needs k = 0 there")
        shape = self.grid.shape
        self._inv = {c: numpy.ones(shape, dtype=numpy.float32) for c in _E}
        self._offdiag: Dict[str, Dict[str, numpy.ndarray]] = {}
        self._conductivity: Dict[str, numpy.ndarray] = {}
        self._absorber_layers: List[AbsorbingFace] = []
        self._poles: List[Tuple[Susceptibility, Dict[str, object]]] = []
        self._pml_spec = None
        self._pml_kwargs = {}
        self._sources: List[Tuple[DepositBox, object]] = []
        self.flux_monitors: List[FluxMonitor] = []
        self.dft_monitors: List[DftFieldsMonitor] = []
        self._default_dft_regions: List[DftFieldsMonitor] = []
        self._automatic_monitors: list = []  # decimation 0: resolved at the freeze
        self.max_steps = int(max_steps)
        self.sim: Optional[HipSimulation] = None
        self._fast_path = None
        self._planned = 0
        self._source_slots: List[List[int]] = []  # engine source indices of each source
        self._shaped = False  # a volume, source or monitor placed on the grid's cells
        if boundaries is not None:
            self.set_boundaries(boundaries)

    def _make_grid(self) -> Grid:
        return Grid(resolution=self._resolution, cell_size=self._cell, courant=self._courant,
                    dimensions=self._dimensions, symmetry=self.symmetry,
                    metallic=tuple(self._boundaries.get(a) == "metallic" for a in range(3)))

    @property
    def is_complex(self) -> bool:
        return self.force_complex_fields

    @property
    def bloch_phases(self) -> Tuple[Optional[complex], ...]:
        """Per axis the Bloch factor (``bloch_phase``) of the full cell, or None."""
        return tuple(bloch_phase(self.k_point[a], self.grid.shape_full[a],
                                 self.grid.resolution) for a in range(3))

    @property
    def sims(self):
        """The parts a batch steps: both parts of complex storage, else None."""
        return getattr(self.sim, "sims", None)

    def _require_unfolded(self, what: str):
        if any(f is not None for f in self.grid.folds):
            raise ValueError(f"{what} on a mirror-folded grid is not built: give the region "
                             "a center and size (the whole cell's images are not placed)")

    # --- configuration (before the first step) ------------------------------------------------
    def _require_unfrozen(self, what: str):
        if self.sim is not None:
            raise RuntimeError(f"{what} must be called before the first step")

    def set_epsilon_components(self, epsilon: Dict[str, object],
                               chi1inv_offdiagonal: Optional[Dict[str, Dict[str, object]]] = None):
        self._require_unfrozen("set_epsilon_components")
        self._shaped = True
        if set(epsilon) != set(_E):
            raise ValueError("epsilon must map exactly Ex, Ey, Ez")
        for c in _E:
            eps = numpy.broadcast_to(numpy.asarray(epsilon[c], dtype=numpy.float64),
                                     self.grid.shape)
            if not numpy.all(numpy.isfinite(eps)) or numpy.any(eps <= 0):
                raise ValueError(f"epsilon[{c}] must be finite and positive")
            # chi1inv is the reciprocal formed in double and rounded once (driver.py:1305)
            self._inv[c] = numpy.ascontiguousarray((1.0 / eps).astype(numpy.float32))
        self._offdiag = {}
        for row, partners in (chi1inv_offdiagonal or {}).items():
            for partner, vol in (partners or {}).items():
                arr = numpy.ascontiguousarray(numpy.broadcast_to(
                    numpy.asarray(vol, dtype=numpy.float32), self.grid.shape))
                if numpy.any(arr):
                    self._offdiag.setdefault(row, {})[partner] = arr

    def set_conductivity(self, conductivity: Dict[str, object]):
        """MEEP's D conductivity (``mp.Medium(D_conductivity=...)``; reference
        ``Fields.set_d_conductivity``, fields.py:775): sigma per component "Dx".."Dz", a
        volume or a scalar; an all-zero one is none."""
        self._set_conductivity_side("D", conductivity)

    def set_b_conductivity(self, conductivity: Dict[str, object]):
        """MEEP's B conductivity, the magnetic half of an ``mp.Absorber`` (reference
        ``Fields.set_b_conductivity``, fields.py:809): sigma per component "Bx".."Bz"."""
        self._set_conductivity_side("B", conductivity)

    def _set_conductivity_side(self, side: str, conductivity: Dict[str, object]):
        self._require_unfrozen("set_conductivity")
        self._shaped = True
        for name in [n for n in self._conductivity if n[0] == side]:
            del self._conductivity[name]
        for name, value in (conductivity or {}).items():
            if len(name) != 2 or name[0] != side or name[1] not in "xyz":
                raise ValueError(f"{side} conductivity components are {side}x, {side}y, {side}z; "
                                 f"got {name!r}")
            # kept in double until it is added to an absorber's and rounded once at the freeze
            vol = numpy.ascontiguousarray(numpy.broadcast_to(
                numpy.asarray(value, dtype=numpy.float64), self.grid.shape))
            if not numpy.all(numpy.isfinite(vol)):
                raise ValueError(f"{name} conductivity must be finite")
            if numpy.any(vol):
                self._conductivity[name] = vol

    def set_boundaries(self, boundaries: Dict[str, str]):
        """Per-axis boundaries, ``{"x": "metallic", ...}`` (the reference driver's
        ``boundaries``): "periodic" (the default) or "metallic" -- perfect electric conductor
        walls on both faces of the axis (stepping.py:140-146). A PML or an absorber on a
        metallic axis grades from the walls without a wrap; an off-diagonal chi1inv row reads
        zero past a wall and couples nothing on its planes (stepping.py:1594-1745)."""
        self._require_unfrozen("set_boundaries")
        kinds = {}
        for axis, kind in _boundary_kinds(boundaries).items():
            key = "xyz"[axis]
            if kind not in ("periodic", "metallic"):
                raise ValueError(f"boundary {key!r} must be 'periodic' or 'metallic', got {kind!r}")
            if kind == "metallic" and self.grid.is_invariant(axis):
                raise ValueError(f"axis {key!r} is invariant: it has no wall")
            if kind == "metallic" and self.k_point[axis] != 0.0:
                raise ValueError(f"k_point is nonzero on {key!r}: a metallic wall gives the axis "
                                 "no lattice vector for a Bloch phase")
            kinds[axis] = kind
        previous = self._boundaries
        self._boundaries = kinds
        grid = self._make_grid()
        if grid.shape != self.grid.shape:
            # a folded axis stores one cell fewer between walls than over a periodic wrap
            if self._shaped:
                self._boundaries = previous
                raise ValueError("on a mirror-folded axis the boundaries set the stored cell "
                                 "count: call set_boundaries before any material, source or "
                                 "monitor")
            self._inv = {c: numpy.ones(grid.shape, dtype=numpy.float32) for c in _E}
        self.grid = grid

    @property
    def wraps(self) -> Tuple[bool, bool, bool]:
        """Per axis: periodic (True) or terminated by metallic walls (reference
        grid.py:1267-1289)."""
        return tuple(self._boundaries.get(a, "periodic") != "metallic" for a in range(3))

    def set_absorber(self, thickness: float, direction: Optional[int] = None,
                     side: Optional[str] = None, R_asymptotic: float = 1e-15):
        """``mp.Absorber(thickness, direction, side, R_asymptotic)``: a graded D and B
        conductivity inside the face(s) (reference absorber.py; MEEP meepgeom.cpp:744-763,
        :1596-1625), added to any material conductivity in double and rounded once. Every
        resolved axis when ``direction`` is None, both sides when ``side`` is None."""
        self._require_unfrozen("set_absorber")
        self._shaped = True
        self._absorber_layers += faces_for(thickness, direction, side, R_asymptotic,
                                            [not self.grid.is_invariant(a) for a in range(3)])

    def conductivity_volumes(self) -> Dict[str, numpy.ndarray]:
        """The float32 sigma each conductive component is stepped with: the material's plus
        the absorber's, summed in double and rounded once."""
        total = {n: v.copy() for n, v in self._conductivity.items()}
        for name, vol in graded_sigma(self.grid, self._absorber_layers,
                                               self.wraps).items():
            total[name] = total[name] + vol if name in total else vol
        return {n: v.astype(numpy.float32) for n, v in total.items() if numpy.any(v)}

    def set_inverse_epsilon(self, inv: Dict[str, numpy.ndarray]):
        self._require_unfrozen("set_inverse_epsilon")
        self._shaped = True
        for c in _E:
            self._inv[c] = numpy.ascontiguousarray(inv[c], dtype=numpy.float32)

    def add_susceptibility(self, term: Susceptibility, sigma):
        self._require_unfrozen("add_susceptibility")
        self._shaped = True
        if not term.stable(self.grid.dt):
            raise ValueError(f"{term} is unstable at dt={self.grid.dt}")
        if not isinstance(sigma, dict):
            sigma = {c: sigma for c in _E}
        self._poles.append((term, dict(sigma)))

    def setup_pml(self, thickness_cells, order: int = 2, R_asymptotic: float = 1e-15):
        self._require_unfrozen("setup_pml")
        probe = PML(self.grid, thickness_cells, order=order, R_asymptotic=R_asymptotic)
        if not probe.is_active:
            raise ValueError("a PML request that absorbs nowhere")
        for a in range(3):
            if self.k_point[a] != 0.0 and any(float(t) > 0 for t in probe.faces[a]):
                raise ValueError(f"k_point is nonzero on {'xyz'[a]!r}, which the PML would "
                                 "terminate: a Bloch phase needs the axis's lattice vector "
                                 "(reference driver.py:2203-2209)")
        self._shaped = True
        self._pml_spec = thickness_cells
        self._pml_kwargs = dict(order=order, R_asymptotic=R_asymptotic)

    def add_source(self, box, time_function):
        """A deposit box (or a list of them, deposited in order) driven by ``time_function``."""
        self._require_unfrozen("add_source")
        boxes = [box] if isinstance(box, DepositBox) else list(box)
        for b in boxes:
            for a, phase in enumerate(self.bloch_phases):
                # stored row 0 of a Yee-shift-0 component on a phased axis holds the window
                # top, the point one lattice vector above the bottom face (reference
                # grid.py:1054-1075): a deposit there would need the phase of its image
                comp = "xyz".index(b.component[1])
                shift0 = (a == comp) if b.magnetic else (a != comp)
                if phase is not None and shift0 and b.lo[a] == 0:
                    raise ValueError(f"a {b.component} source on stored row 0 of the Bloch-phased "
                                     f"axis {'xyz'[a]} is not built: move it off the cell face")
        self._shaped = True
        for b in boxes:
            self._sources.append((b, time_function))

    def add_volume_source(self, component: str, center: Sequence[float], size: Sequence[float],
                          time_function, amplitude: complex = 1.0,
                          is_integrated: bool = False) -> List[DepositBox]:
        """``mp.Source(src, component, center, size, amplitude, is_integrated)``: placed as MEEP
        places it on this grid and its boundaries (``meep_volume_source``); returns the boxes."""
        self._require_unfrozen("add_volume_source")
        for a, phase in enumerate(self.bloch_phases):
            half = 0.5 * self.grid.shape_full[a] * self.grid.dx
            lo, hi = center[a] - 0.5 * size[a], center[a] + 0.5 * size[a]
            if phase is not None and (lo < -half + self.grid.dx or hi > half - self.grid.dx):
                raise ValueError(f"a source reaching the Bloch-phased boundary of "
                                 f"{'xyz'[a]} is not built: its wrapped part would carry the "
                                 "phase")
        boxes = meep_volume_source(self.grid, component, center, size, amplitude,
                                   metallic=tuple(not w for w in self.wraps),
                                   integrated=is_integrated)
        self.add_source(boxes, time_function)
        return boxes

    def add_flux_monitor(self, normal: int, lo: Sequence[int], cnt: Sequence[int],
                         fcen: float = None, df: float = None, nfreq: int = None,
                         frequencies: Sequence[float] = None, decimation: int = 1) -> FluxMonitor:
        self._require_unfrozen("add_flux_monitor")
        self._shaped = True
        freqs = (numpy.asarray(frequencies, dtype=numpy.float64) if frequencies is not None
                 else meep_frequency_span(fcen, df, nfreq))
        mon = FluxMonitor(int(normal), tuple(lo), tuple(cnt), freqs, self.grid.dx, self.grid.dt,
                          decimation or 1)
        self._note_automatic(mon, decimation)
        self.flux_monitors.append(mon)
        return mon

    def add_flux_region(self, center: Sequence[float], size: Sequence[float],
                        fcen: float = None, df: float = None, nfreq: int = None,
                        frequencies: Sequence[float] = None, decimation: int = 1) -> FluxMonitor:
        """``sim.add_flux(fcen, df, nfreq, mp.FluxRegion(center, size))``: the plane placed and
        weighted as MEEP places it (coordinates with their origin at the cell centre).
        ``decimation`` is the accumulation stride MEEP resolved for the monitor."""
        self._require_unfrozen("add_flux_region")
        self._shaped = True
        freqs = (numpy.asarray(frequencies, dtype=numpy.float64) if frequencies is not None
                 else meep_frequency_span(fcen, df, nfreq))
        mon = FluxMonitor.from_region(self.grid, tuple(center), tuple(size), freqs,
                                      int(decimation) or 1)
        self._note_automatic(mon, decimation)
        self.flux_monitors.append(mon)
        return mon

    def add_dft_monitor(self, frequencies=None, components: Sequence[str] = ("Ex", "Ey", "Ez"),
                        center: Optional[Sequence[float]] = None,
                        size: Optional[Sequence[float]] = None, *, fcen: float = None,
                        df: float = None, nfreq: int = None, decimation: int = 1,
                        yee_grid: bool = False) -> DftFieldsMonitor:
        """``sim.add_dft_fields(components, fcen, df, nfreq, center=, size=)`` (reference
        ``FdtdDriver.add_dft_monitor``, driver.py:2783): the DFT of ``components`` over the
        volume (center, size), or, without one, over the absorber's interior (the whole grid
        without an absorber), resolved when stepping starts. ``decimation`` is an explicit
        positive accumulation stride."""
        self._require_unfrozen("add_dft_monitor")
        if center is None:
            self._require_unfolded("a DFT monitor without a region")
        self._shaped = True
        freqs = (numpy.atleast_1d(numpy.asarray(frequencies, dtype=numpy.float64))
                 if frequencies is not None else meep_frequency_span(fcen, df, nfreq))
        if (center is None) != (size is None):
            raise ValueError("a DFT monitor region needs both center and size, or neither")
        if size is not None and any(float(v) < 0 for v in size):
            raise ValueError(f"monitor size must be non-negative, got {tuple(size)}")
        if center is not None:
            mon = DftFieldsMonitor.from_region(self.grid, tuple(components), tuple(center),
                                               tuple(size), freqs, int(decimation) or 1,
                                               yee_grid)
        else:  # placed at freeze, when the absorber is known
            mon = DftFieldsMonitor(tuple(components), (0, 0, 0), tuple(self.grid.shape), freqs,
                                   self.grid.dt, int(decimation) or 1, yee_grid,
                                   invariant=tuple(self.grid.is_invariant(a) for a in range(3)))
            self._default_dft_regions.append(mon)
        self._note_automatic(mon, decimation)
        self.dft_monitors.append(mon)
        return mon

    def _note_automatic(self, mon, decimation) -> None:
        if int(decimation) < 0:
            raise ValueError(f"decimation must be >= 0 (0: automatic), got {decimation!r}")
        if int(decimation) == 0:
            self._automatic_monitors.append(mon)

    def automatic_decimation(self, frequencies: Sequence[float]) -> int:
        """MEEP's automatic decimation (``decimation_factor == 0``; reference
        ``_automatic_decimation_factor`` driver.py:3101-3163, MEEP dft.cpp:195-216):
        ``floor(1 / (dt * (freq_max + src_freq_max)))``, at least 1, where each Gaussian
        source contributes ``|frequency| + fwidth/2`` and continuous or custom sources nothing
        (1 when no Gaussian source or no nonzero monitor frequency). Resolved at the freeze,
        when every source is known."""
        src_max = 0.0
        for _, fn in self._sources:
            if isinstance(fn, GaussianSource):
                src_max = max(src_max, abs(float(fn.frequency)) + 0.5 * float(fn.fwidth))
        mon_max = max((abs(float(f)) for f in numpy.atleast_1d(frequencies)), default=0.0)
        if mon_max <= 0.0 or src_max <= 0.0:
            return 1
        return max(1, int(math.floor(1.0 / (self.grid.dt * (mon_max + src_max)))))

    def _place_default_dft_regions(self, pml) -> None:
        """The absorber's interior: each face trimmed by the whole cells it grades."""
        for mon in self._default_dft_regions:
            lo, cnt = [0, 0, 0], list(self.grid.shape)
            if pml is not None:
                for a in range(3):
                    low, high = (int(math.ceil(float(t) - 1e-9)) for t in pml.faces[a])
                    lo[a], cnt[a] = low, self.grid.shape[a] - low - high
                    if cnt[a] < 1:
                        raise ValueError(f"axis {a}: the absorber leaves no interior")
            mon.lo, mon.cnt = tuple(lo), tuple(cnt)
        self._default_dft_regions = []

    @classmethod
    def from_workload(cls, wl: Workload) -> "FdtdDriver":
        drv = cls.__new__(cls)
        drv.grid = wl.grid
        drv._cell, drv._resolution = tuple(wl.grid.cell_size), wl.grid.resolution
        drv._courant, drv._dimensions = wl.grid.courant, wl.grid.dimensions
        drv.force_complex_fields = False
        drv.symmetry = _mirror_entries(wl.grid.symmetry)
        drv.k_point = (0.0, 0.0, 0.0)
        drv._shaped = True
        drv._inv = dict(wl.inv)
        drv._offdiag = {}
        drv._conductivity = {}
        drv._absorber_layers = []
        drv._boundaries = {}
        drv._poles = list(wl.poles)
        drv._pml_spec = wl.pml_faces
        drv._pml_kwargs = {}
        drv._sources = list(wl.sources)
        drv.flux_monitors = []
        drv.dft_monitors = []
        drv._default_dft_regions = []
        drv._automatic_monitors = []
        drv.max_steps = max(wl.steps, 1)
        drv.sim = None
        drv._fast_path = None
        drv._planned = 0
        for p in wl.planes:
            drv.add_flux_monitor(p.normal, p.lo, p.cnt, frequencies=p.frequencies)
        return drv

    # --- the run -------------------------------------------------------------------------------
    @property
    def step_count(self) -> int:
        return 0 if self.sim is None else self.sim.step_count

    @property
    def time(self) -> float:
        return self.step_count * self.grid.dt

    def freeze(self, planned_steps: Optional[int] = None):
        """Create the device simulation and upload every table (the reference's first-step
        freeze). On complex storage: two parts, linked by the Bloch connections."""
        if self.sim is not None:
            return self.sim
        steps = int(planned_steps or self.max_steps)
        pml = (PML(self.grid, self._pml_spec, wraps=self.wraps, **self._pml_kwargs)
               if self._pml_spec else None)
        for mon in self._automatic_monitors:
            mon.decimation = self.automatic_decimation(mon.frequencies)
        self._place_default_dft_regions(pml)
        parts = []
        try:
            for part in range(2 if self.force_complex_fields else 1):
                parts.append(self._build_part(pml, steps, part))
            if self.force_complex_fields:
                from .capsule import bloch_connection
                from .engine import ComplexSimulation, link_parts

                phases = self.bloch_phases
                kinds = [0 if z is None else bloch_connection(z) for z in phases]
                link_parts(parts[0], parts[1], kinds,
                           [(1.0, 0.0) if z is None else (z.real, z.imag) for z in phases])
                sim = ComplexSimulation(parts[0], parts[1])
            else:
                sim = parts[0]
            for mon in self.flux_monitors:
                mon.attach(sim)
            for mon in self.dft_monitors:
                mon.attach(sim)
        except Exception:
            for p in parts:
                p.close()
            raise
        plan = plan_fast_path(sim, pml, self.grid, self._sources)
        if plan is None:
            sim.close()
            raise FastPathError(f"no engine plan: {last_dispatch_report()['reasons']}")
        self.sim = sim
        self._fast_path = plan
        self._planned = steps
        return sim

    def _build_part(self, pml, steps: int, part: int) -> HipSimulation:
        """One part's device simulation with every table; part 1 (the imaginary part) deposits
        each source's imaginary line, the real line of amplitudes ``(amp_im, -amp_re)``
        (reference sources.py:1357-1482)."""
        sim = HipSimulation(self.grid, pml)
        try:
            for axis, kind in self._boundaries.items():
                if self.grid.folds[axis] is None:  # a fold's walls are part of its storage
                    sim.set_boundary(axis, kind)
            for c in _E:
                sim.set_inverse_epsilon(c, self._inv[c])
            for row, partners in self._offdiag.items():
                for partner, vol in partners.items():
                    sim.set_offdiagonal(row, partner, vol)
            for term, sigma in self._poles:
                sim.add_pole(term, sigma)
            if self._conductivity or self._absorber_layers:
                if any(f is not None for f in self.grid.folds):
                    raise ValueError("a conductivity on a mirror-folded grid is not built: "
                                     "MEEP's chunk division of the halved cell is not restated")
                division = ChunkDivision.one_rank(
                    self.grid, pml.faces if pml is not None else ((0, 0), (0, 0), (0, 0)),
                    self.wraps)
                for name, vol in self.conductivity_volumes().items():
                    # MEEP keeps a chunk's conductivity only where it samples a nonzero sigma
                    sim.set_conductivity(name, vol, division.conductive_mask(vol, IYEE[name]))
            # a polarization box goes in as its pieces (off ghost and wall rows, plus its
            # images on the ghost rows); each source's pieces share its current table
            slots, count = [], 0
            metallic = tuple(not w for w in self.wraps)
            for box, fn in self._sources:
                table = current_table(fn, self.grid.dt, steps, box.magnetic, box.integrated)
                axis = "xyz".index(box.component[1])
                mine = []
                for piece in polarization_boxes(self.grid, box, metallic):
                    amp_re, amp_im = piece.amp_re, piece.amp_im
                    if part == 1:
                        amp_re, amp_im = (numpy.asarray(piece.amp_im, dtype=numpy.float64),
                                          -numpy.asarray(piece.amp_re, dtype=numpy.float64))
                    mine.append(count)
                    count += 1
                    sim.add_source(piece.magnetic, axis, piece.lo, piece.cnt, amp_re, amp_im,
                                   table, self.grid.dt, piece.integrated)
                slots.append(mine)
            self._source_slots = slots
            if pml is None and not self._poles and not self._offdiag:
                sim.prime_e()
        except Exception:
            sim.close()
            raise
        return sim

    #: The consults of one step in the reference driver's order (driver.py:3505-3596).
    CONSULTS = ("step_B", "fill_B", "fill_folded_far_ghosts_B", "update_H", "step_D", "fill_D",
                "fill_folded_far_ghosts_D", "update_E", "update_P")

    def _extend_sources(self, steps: int):
        """Source tables long enough for ``steps`` steps from step 0: a run past the planned
        length extends them on the device (``fdtd_set_source_currents``), doubling the plan,
        where the reference evaluates each current as it goes (sources.py:1844-1936)."""
        if steps <= self._planned:
            return
        planned = max(int(steps), 2 * self._planned)
        for slots, (box, fn) in zip(self._source_slots, self._sources):
            table = current_table(fn, self.grid.dt, planned, box.magnetic, box.integrated)
            for q in slots:
                self.sim.set_source_currents(q, table)
        self._planned = planned

    def step(self):
        """One time step through the dispatch seam: each consult must be answered by the plan
        (there is no array path to fall back on), the fused launches at step_B and step_D."""
        sim = self.freeze()
        self._extend_sources(sim.step_count + 1)
        fast = self._fast_path
        for slot in self.CONSULTS:
            if not fast.dispatch(slot, sim):
                raise FastPathError(f"the engine plan declined {slot}")

    def run(self, steps: Optional[int] = None, until: Optional[float] = None):
        if (steps is None) == (until is None):
            raise ValueError("give exactly one of steps or until")
        if until is not None:
            steps = int(math.ceil(until / self.grid.dt - 1e-9))
        sim = self.freeze(max(int(steps) + self.step_count, self.max_steps))
        self._extend_sources(sim.step_count + int(steps))
        sim.run(int(steps))

    def synchronize(self):
        if self.sim is not None:
            self.sim.synchronize()

    def get_array(self, component: str) -> numpy.ndarray:
        """A stored volume in the reference's layout (C-order (nx, ny, nz)): float32, or
        complex64 on complex storage (the two parts' words, side by side)."""
        sim = self.freeze()
        if component[0] == "H" and sim.pml is None:
            component = "B" + component[1]
        if self.force_complex_fields:
            return sim.download_complex(component)
        return sim.download(component)

    def get_meep_array(self, component: str, center=None, size=None) -> numpy.ndarray:
        """MEEP's ``sim.get_array(component)`` over the whole cell or a ``center``/``size``
        box: the centred lattice, word for word (``readback.meep_array``), where ``get_array``
        gives the stored words. Complex64 on complex storage (with the Bloch window's
        phases); a mirror-folded grid's images unfolded over the whole cell with their
        parities."""
        sim = self.freeze()
        pml = sim.pml
        division = ChunkDivision.one_rank(
            self.grid, pml.faces if pml is not None else ((0, 0), (0, 0), (0, 0)), self.wraps)
        return meep_array(self.get_array, component, self.grid, self.wraps, division,
                          pml is not None, sim.step_count, center=center, size=size,
                          phases=self.bloch_phases)

    def fast_path_report(self) -> Dict[str, object]:
        from .engine import device_info

        rep = last_dispatch_report()
        rep.update({"engine": "hip", "device": device_info(),
                    "library": "meep_gpu/_lib/libmeep_fdtd.so",
                    "launches": self._fast_path.launches[0] if self._fast_path else 0})
        return rep

    def close(self):
        if self.sim is not None:
            self.sim.close()
            self.sim = None
