"""Synthetic FDTD workloads: the sphere and slab cases the benchmark and the tests step.

A workload is a plain description -- grid, absorber faces, diagonal inverse permittivity per
component, poles, deposit boxes with their time functions, flux planes -- that either stepper
(the HIP engine or the CPU restatement) can be built from, so both step identical inputs.

* ``sphere``: a dielectric sphere (eps inside, vacuum outside) centred in a cubic periodic cell
  under a UPML on all six faces, an Ez Gaussian point source, x-normal or box-face flux planes.
  The chi1inv of each E component is sampled at that component's own Yee position with a
  sub-pixel fill fraction ``f`` (from the signed distance to the surface) as
  ``chi1inv = f/eps + (1 - f)/1`` (diagonal only).
* ``slab``: a Lorentz-dispersive slab, normal x, spanning the transverse cell, a planar Ez
  Gaussian source, a reflection and a transmission plane.
"""

from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Sequence, Tuple

import numpy

from .dispersion import LORENTZIAN, Susceptibility
from .grid import Grid
from .sources import DepositBox, GaussianSource, meep_volume_source

_E = ("Ex", "Ey", "Ez")
#: Yee offsets in half cells per component (reference fields.py:218-223).
IYEE = {"Ex": (1, 0, 0), "Ey": (0, 1, 0), "Ez": (0, 0, 1),
        "Hx": (0, 1, 1), "Hy": (1, 0, 1), "Hz": (1, 1, 0)}


@dataclass
class FluxPlane:
    normal: int
    lo: Tuple[int, int, int]
    cnt: Tuple[int, int, int]
    frequencies: numpy.ndarray

    @property
    def cells(self) -> int:
        return int(numpy.prod(self.cnt))


@dataclass
class Workload:
    name: str
    grid: Grid
    pml_faces: object
    inv: Dict[str, numpy.ndarray]
    poles: List[Tuple[Susceptibility, Dict[str, object]]] = field(default_factory=list)
    sources: List[Tuple[DepositBox, object]] = field(default_factory=list)
    planes: List[FluxPlane] = field(default_factory=list)
    steps: int = 0

    @property
    def cells(self) -> int:
        return self.grid.total_cells


def meep_frequency_span(fcen: float, df: float, nfreq: int) -> numpy.ndarray:
    """MEEP's (fcen, df, nfreq): nfreq points over fcen +- df/2; one point is fcen (dft.py:137)."""
    if nfreq == 1:
        return numpy.array([float(fcen)])
    return numpy.linspace(fcen - 0.5 * df, fcen + 0.5 * df, nfreq)


def _yee_coordinates(grid: Grid, component: str):
    """Physical coordinates (origin at the cell centre) of every sample of a component."""
    shift = IYEE[component]
    axes = []
    for axis, n in enumerate(grid.shape):
        axes.append(-0.5 * n * grid.dx + (numpy.arange(n) + 0.5 * shift[axis]) * grid.dx)
    return axes


def _sphere_inverse_epsilon(grid: Grid, radius: float, eps_in: float, eps_out: float = 1.0
                            ) -> Dict[str, numpy.ndarray]:
    """Sub-pixel chi1inv of a sphere: the fill fraction of each Yee cell from the signed
    distance to the surface, f = clip(1/2 - d/dx, 0, 1) (a locally planar interface through
    the cell), then chi1inv = f/eps_in + (1 - f)/eps_out."""
    # f is exactly 1 (0) more than half a cell inside (outside) the surface, so only a band of
    # +-1 cell around it takes the formula; the rest are the two constants the formula gives
    # there, element for element the same words (the admission of a served solve waits on this)
    inside = numpy.float32(1.0 / eps_in + (1.0 - 1.0) / eps_out)
    outside = numpy.float32(0.0 / eps_in + (1.0 - 0.0) / eps_out)
    near2 = (radius - grid.dx) ** 2 if radius > grid.dx else -1.0
    far2 = (radius + grid.dx) ** 2
    out = {}
    for c in _E:
        x, y, z = _yee_coordinates(grid, c)
        r2 = (x[:, None, None] ** 2 + y[None, :, None] ** 2) + z[None, None, :] ** 2
        vol = numpy.where(r2 <= near2, inside, outside).astype(numpy.float32)
        band = (r2 > near2) & (r2 < far2)
        r = numpy.sqrt(r2[band])
        f = numpy.clip(0.5 - (r - radius) / grid.dx, 0.0, 1.0)
        vol[band] = (f / eps_in + (1.0 - f) / eps_out).astype(numpy.float32)
        out[c] = vol
    return out


def _ez_point(grid: Grid, x: float, y: float, z: float) -> List[DepositBox]:
    """An Ez point source at (x, y, z), deposited as MEEP deposits a zero-size source."""
    return meep_volume_source(grid, "Ez", (x, y, z), (0.0, 0.0, 0.0))


def outside_absorber(grid: Grid, pml_cells: int) -> numpy.ndarray:
    """Cells at least ``pml_cells`` from every face of a resolved axis.

    A Lorentz susceptibility must stay out of the UPML: with the pole's sigma inside the layer
    the coupled update grows without bound (the oracle, which restates the reference's array
    path, blows up the same way on those inputs), so workloads keep their dispersive material
    inside this mask.
    """
    mask = numpy.ones(grid.shape, dtype=bool)
    for axis, n in enumerate(grid.shape):
        if grid.is_invariant(axis) or pml_cells <= 0:
            continue
        idx = numpy.arange(n)
        keep = (idx >= pml_cells) & (idx < n - pml_cells)
        shape = [1, 1, 1]
        shape[axis] = n
        mask &= keep.reshape(shape)
    return mask


def _plane_index(x: float, edge: int, dx: float) -> int:
    """The cell whose centre is nearest the coordinate ``x`` (origin at the cell centre): flux
    planes sample cell-centred fields."""
    return int(math.floor((x + 0.5 * edge * dx) / dx))


def sphere(edge: int = 80, resolution: float = 10.0, pml_cells: int = 8,
           radius_cells: float = 16.0, eps: float = 2.25, fcen: float = 0.15, df: float = 0.1,
           nfreq: int = 50, planes: str = "one", steps: int = 2000,
           source_x: float = -2.0, plane_x: float = 2.0) -> Workload:
    """The dielectric sphere under a UPML (configs 1 and 2 of the benchmark table)."""
    L = edge / resolution
    grid = Grid(resolution=resolution, cell_size=(L, L, L))
    assert grid.shape == (edge, edge, edge), grid.shape
    r = radius_cells / resolution
    inv = _sphere_inverse_epsilon(grid, r, eps)
    freqs = meep_frequency_span(fcen, df, nfreq)
    src = GaussianSource(fcen, df)
    wl = Workload(f"sphere_{edge}", grid, pml_cells, inv, steps=steps)
    wl.sources.extend((box, src) for box in _ez_point(grid, source_x, 0.0, 0.0))
    c = edge // 2
    if planes == "one":
        half = edge // 4
        i = _plane_index(plane_x, edge, grid.dx)
        if not pml_cells <= i < edge - pml_cells:
            raise ValueError(f"flux plane at x={plane_x} falls outside the interior")
        wl.planes.append(FluxPlane(0, (i, c - half, c - half), (1, 2 * half, 2 * half), freqs))
    elif planes == "box8":
        b = int(round(radius_cells * 1.25))
        lo_b, hi_b = c - b, c + b
        for axis in range(3):
            for pos in (lo_b, hi_b):
                lo = [lo_b, lo_b, lo_b]
                cnt = [2 * b, 2 * b, 2 * b]
                lo[axis], cnt[axis] = pos, 1
                wl.planes.append(FluxPlane(axis, tuple(lo), tuple(cnt), freqs))
        inner = edge - 2 * pml_cells
        for pos in (pml_cells + 4, edge - pml_cells - 4):
            wl.planes.append(FluxPlane(0, (pos, pml_cells, pml_cells), (1, inner, inner), freqs))
    elif planes != "none":
        raise ValueError(f"planes must be 'one', 'box8' or 'none', got {planes!r}")
    return wl


def slab(edge: int = 160, resolution: float = 20.0, pml_cells: int = 10, thickness: int = 20,
         eps_inf: float = 2.0, sigma: float = 1.5, f0: float = 1.1, gamma: float = 0.05,
         fcen: float = 0.8, df: float = 0.6, nfreq: int = 200, steps: int = 20000) -> Workload:
    """The Lorentz slab (config 3): normal x, centred, full transverse extent."""
    L = edge / resolution
    grid = Grid(resolution=resolution, cell_size=(L, L, L))
    lo_x = edge // 2 - thickness // 2
    mask = numpy.zeros(grid.shape, dtype=bool)
    mask[lo_x:lo_x + thickness, :, :] = True
    inv = {}
    for c in _E:
        inv[c] = numpy.where(mask, numpy.float32(1.0 / eps_inf), numpy.float32(1.0)).astype(
            numpy.float32)
    sig = numpy.where(mask, numpy.float32(sigma), numpy.float32(0.0)).astype(numpy.float32)
    wl = Workload(f"slab_{edge}", grid, pml_cells, inv, steps=steps)
    wl.poles.append((Susceptibility(f0, gamma, LORENTZIAN), {c: sig for c in _E}))
    src_x = pml_cells + 5
    x = -0.5 * L + (src_x + 0.25) * grid.dx  # a quarter cell past the sample: both weights live
    pulse = GaussianSource(fcen, df)
    wl.sources.extend((box, pulse)
                      for box in meep_volume_source(grid, "Ez", (x, 0.0, 0.0), (0.0, L, L)))
    freqs = meep_frequency_span(fcen, df, nfreq)
    inner = edge - 2 * pml_cells
    for pos in (src_x + 5, lo_x + thickness + 10):
        wl.planes.append(FluxPlane(0, (pos, pml_cells, pml_cells), (1, inner, inner), freqs))
    return wl


def config(name: str, **overrides) -> Workload:
    """The named benchmark configurations, with optional size overrides for tests."""
    if name == "sphere80":
        return sphere(**{**dict(edge=80, resolution=10.0, radius_cells=16.0, nfreq=50,
                                planes="one", steps=2000), **overrides})
    if name == "sphere192":
        return sphere(**{**dict(edge=192, resolution=24.0, radius_cells=38.4, nfreq=100,
                                planes="box8", steps=10000), **overrides})
    if name == "slab160":
        return slab(**overrides)
    raise KeyError(name)
