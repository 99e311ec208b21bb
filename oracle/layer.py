"""Setup-time restatements the oracle steps with: grid size, UPML table, chunk lines, ADE words.

ORACLE / test infrastructure only. Each function restates one reference routine on the
periodic (wrapping) Cartesian grids the fixture cases and the benchmark workloads use:

* ``cell_count``        <- ``grid.meep_cell_count``               grid.py:288-289
* ``half_cell_extent``  <- ``pml.half_cell_extent``               pml.py:132-152
* ``graded_sigma``      <- ``PML._graded_sigma``                  pml.py:469-598 (fill loop at
                           its end), ``PML.prefac`` pml.py:436-467
* ``axis_coefficients`` <- ``PML._compute_sigma_profile``         pml.py:633-671
* ``membership_line``   <- ``chunk_map.one_rank_line``            chunk_map.py:180-233
* ``ade_constants``     <- ``Susceptibility.realnum_constants``   dispersion.py:243-290

Pinned by: the fixture digests (every case with a layer reaches both curl forms and both
constitutive forms), the MEEP update_P words of tests/golden/ade_words.json, and the closed
form of sigma at the wall (tested in tests/test_oracle_layer.py).
"""

from __future__ import annotations

import math
from typing import Optional, Tuple

import numpy

f32 = numpy.float32


def cell_count(length: float, resolution: float) -> int:
    return int(length * resolution + 0.5)


def half_cell_extent(cells: float) -> int:
    return int(2.0 * float(cells) + 0.5)


def prefac(cells: float, dx: float, order: int = 2, ln_reflection: Optional[float] = None,
           reflection: float = 1e-15) -> float:
    """``-ln(R) / (4 * dx_pml * 1/(order+1))``; ``ln_reflection`` replaces ``-ln(R)`` when given."""
    dx_pml = cells * dx
    if dx_pml <= 0:
        return 0.0
    profile_integral = 1.0 / (order + 1)
    numerator = ln_reflection if ln_reflection is not None else float(-numpy.log(reflection))
    return float(numerator / (4.0 * dx_pml * profile_integral))


def graded_sigma(n_cells: int, faces: Tuple[float, float], offset: float, dt: float, dx: float,
                 order: int = 2, ln_reflection: Optional[float] = None,
                 wraps: bool = True, wall_high: Optional[int] = None) -> numpy.ndarray:
    """float32 ``0.5*dt*prefac*u**order`` on the integer (offset 0) or half (0.5) lattice.

    ``wall_high`` is the upper wall's half-cell index, ``2 * n_cells`` unless given: a folded
    periodic axis stores one row past its window, and its wall is ``2 * owned`` (pml.py:600-611).
    """
    sig = numpy.zeros(n_cells, dtype=numpy.float32)
    half_index = 2 * numpy.arange(n_cells) + int(2 * offset)
    if wraps:
        half_index = (half_index - 1) % (2 * n_cells) + 1
    if wall_high is None:
        wall_high = 2 * n_cells
    for wall, n_pml in ((0, faces[0]), (wall_high, faces[1])):
        if n_pml <= 0:
            continue
        scale = (0.5 * dt) * prefac(n_pml, dx, order, ln_reflection)
        extent = half_cell_extent(n_pml)
        distance = numpy.abs(half_index - wall)
        overhang = extent - 2.0 * n_pml
        u = 1.0 + (overhang - distance) / (2.0 * n_pml)
        inside = distance < extent
        sig[inside] = scale * u[inside] ** order
    return sig


def axis_coefficients(n_cells: int, faces: Tuple[float, float], offset: float, dt: float,
                      dx: float, order: int = 2, ln_reflection: Optional[float] = None,
                      wraps: bool = True, wall_high: Optional[int] = None
                      ) -> Tuple[numpy.ndarray, numpy.ndarray, numpy.ndarray]:
    """``(kms, sinv, kps)`` float32 vectors: ``1-sig``, ``1/(1+sig)``, ``1+sig`` with kappa 1."""
    sig = graded_sigma(n_cells, faces, offset, dt, dx, order, ln_reflection, wraps, wall_high)
    kap = numpy.ones(n_cells, dtype=numpy.float32)
    sinv = (f32(1.0) / (kap + sig)).astype(numpy.float32)
    return (kap - sig).astype(numpy.float32), sinv, (kap + sig).astype(numpy.float32)


def membership_line(n_cells: int, faces: Tuple[float, float], half_integer: bool,
                    wraps: bool = True) -> numpy.ndarray:
    """One-rank chunk membership: does the chunk owning this stored row carry the layer?"""
    low, high = faces
    inside = numpy.zeros(n_cells, dtype=bool)
    if low > 0:
        span = min(n_cells, int(low + 1.5))
        if half_integer:
            inside[0:span] = True
        else:
            inside[1:min(n_cells, span + 1)] = True
    if high > 0:
        span = min(n_cells, int(high + 1.5))
        if half_integer:
            inside[n_cells - span:] = True
        else:
            inside[max(0, n_cells - span + 1):] = True
            if wraps:
                inside[0] = True
    return inside


def ade_constants(frequency: float, gamma: float, drude: bool, dt: float
                  ) -> Tuple[numpy.float32, numpy.float32, numpy.float32, numpy.float32]:
    """MEEP's four realnum words ``(two_minus_od, gamma1, omega0dtsqr, gamma1inv)``."""
    omega_0 = f32(frequency)
    gam = f32(gamma)
    dt_r = f32(dt)
    omega2pi = f32((2.0 * math.pi) * float(omega_0))
    g2pi = f32(float(gam * f32(2)) * math.pi)
    omega0dtsqr = ((omega2pi * omega2pi) * dt_r) * dt_r
    half = (g2pi * dt_r) / f32(2)
    gamma1inv = f32(1) / (f32(1) + half)
    gamma1 = f32(1) - half
    denom = f32(0) if drude else omega0dtsqr
    return f32(2) - denom, gamma1, omega0dtsqr, gamma1inv


def pml_faces(spec, dimensions: int = 3) -> Tuple[Tuple[float, float], ...]:
    """Per-face cell counts from a scalar, ``{axis: t | (lo, hi)}`` or a 3-sequence spec."""
    def pair(v):
        if isinstance(v, (tuple, list)):
            return (v[0], v[1])
        return (v, v)
    if spec is None:
        return ((0, 0), (0, 0), (0, 0))
    if isinstance(spec, dict):
        faces = [(0, 0), (0, 0), (0, 0)]
        for key, value in spec.items():
            faces["xyz".index(key)] = pair(value)
        return tuple(faces)
    if isinstance(spec, (tuple, list)):
        return tuple(pair(v) for v in spec)
    faces = [pair(spec)] * 3
    if dimensions == 2:
        faces[2] = (0, 0)
    return tuple(faces)
