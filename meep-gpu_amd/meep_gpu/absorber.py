"""``mp.Absorber``: a graded scalar D and B conductivity on a face, in place of a PML.

An absorber is a material, not a boundary: the face keeps its (here periodic) wrap and the
layer is a conductivity that turns on gradually inside it, on both the D and the B side (MEEP
python/simulation.py:316; structure.cpp:377-379). The volumes computed here go through the
engine's conductivity path (``fdtd_set_conductivity``), which steps MEEP's conductive curl forms.

The rule, as the reference restates MEEP (absorber.py:1-316; meepgeom.cpp:744-763 and
:1596-1625; structure.cpp:868-895):

* a profile table per layer, sampled at HALF pixels: ``dx = 1/(2 resolution)``,
  ``N = int(L/dx + 0.5)``, ``prefac = -ln(R) / (4 L * 1/3)`` (the integral of the quadratic
  profile u^2 over [0, 1]), ``prof[i] = prefac * (i/N)^2``;
* at a coordinate x, for a high-side layer ``edge = Lcell/2 - L`` and ``u = N (x - edge)/L``
  where ``x >= edge``, for a low-side one ``edge = L - Lcell/2`` and ``u = N (edge - x)/L``
  where ``x <= edge``; then ``i = int(u)`` and a linear interpolation between ``prof[i]`` and
  ``prof[i+1]``, clamped to ``prof[N]`` once ``i >= N``; the layers' contributions add;
* each D and B component is sampled at its own Yee position, formed in float64 from one
  doubled integer ``(-n + 2 i + shift) * dx/2``; on a wrapping axis an integer-lattice
  component's index 0 is MEEP's seam copy of index n, and takes that coordinate; a metallic
  axis does not wrap, and its index 0 is the wall at ``-L/2`` (absorber.py:221-246);
* the sigma stays float64 until it is added to any material conductivity, then is rounded
  to float32 once.
"""

from __future__ import annotations

import math
from dataclasses import dataclass
from typing import Dict, Optional, Sequence, Tuple

import numpy

REFLECTION_DEFAULT = 1e-15
_U_SQUARED_MEAN = 1.0 / 3.0
GRADED_COMPONENTS: Tuple[str, ...] = ("Dx", "Dy", "Dz", "Bx", "By", "Bz")

# Yee shift (1 = half-integer position) of each component along x, y, z
_SHIFTS = {"Dx": (1, 0, 0), "Dy": (0, 1, 0), "Dz": (0, 0, 1),
           "Bx": (0, 1, 1), "By": (1, 0, 1), "Bz": (1, 1, 0)}


@dataclass(frozen=True)
class AbsorbingFace:
    """One absorbing face: ``thickness`` in length units (MEEP's L, not a cell count) on side
    ``"low"`` or ``"high"`` of axis 0/1/2."""
    thickness: float
    axis: int
    side: str
    R_asymptotic: float = REFLECTION_DEFAULT

    def __post_init__(self):
        if self.axis not in (0, 1, 2):
            raise ValueError(f"an absorbing face needs axis 0, 1 or 2 (got {self.axis!r})")
        if self.side not in ("low", "high"):
            raise ValueError(f"an absorbing face is on side 'low' or 'high' (got {self.side!r})")
        if not math.isfinite(self.thickness) or self.thickness <= 0.0:
            raise ValueError(f"an absorbing face needs a positive finite thickness "
                             f"(got {self.thickness!r})")
        if not math.isfinite(self.R_asymptotic) or not 0.0 < self.R_asymptotic < 1.0:
            raise ValueError(f"the asymptotic reflection is a fraction strictly between 0 and 1 "
                             f"(got {self.R_asymptotic!r})")


def faces_for(thickness: float, direction: Optional[int] = None, side: Optional[str] = None,
               R_asymptotic: float = REFLECTION_DEFAULT,
               resolved: Sequence[bool] = (True, True, True)):
    """``mp.Absorber(thickness, direction, side, R_asymptotic)`` as layers: every resolved axis
    when ``direction`` is None, both sides when ``side`` is None."""
    axes = [a for a in range(3) if resolved[a]] if direction is None else [int(direction)]
    sides = ("low", "high") if side is None else (side,)
    return [AbsorbingFace(float(thickness), a, s, float(R_asymptotic)) for a in axes for s in sides]


def profile_table(thickness: float, resolution: float,
                     R_asymptotic: float = REFLECTION_DEFAULT) -> Tuple[int, numpy.ndarray]:
    """(N, prof[0..N]) of one layer: the table sampled at half pixels."""
    if resolution <= 0.0 or not math.isfinite(resolution):
        raise ValueError(f"the resolution must be positive and finite (got {resolution!r})")
    half_pixel = 0.5 / resolution
    count = int(thickness / half_pixel + 0.5)
    if count < 1:
        raise ValueError(f"a face {thickness!r} thick holds {count} half pixels at resolution "
                         f"{resolution!r}: the profile table needs at least one")
    prefac = (-math.log(R_asymptotic)) / (4.0 * thickness * _U_SQUARED_MEAN)
    index = numpy.arange(count + 1, dtype=numpy.float64)
    return count, prefac * (index / count) ** 2


def sigma_on_axis(coordinates: numpy.ndarray, layer: AbsorbingFace, half_length: float,
                     resolution: float) -> numpy.ndarray:
    """One layer's float64 sigma at the given coordinates (zero outside the layer)."""
    count, profile = profile_table(layer.thickness, resolution, layer.R_asymptotic)
    L = float(layer.thickness)
    x = numpy.asarray(coordinates, dtype=numpy.float64)
    if layer.side == "high":
        edge = half_length - L
        inside = x >= edge
        u = count * (x - edge) / L
    else:
        edge = L - half_length
        inside = x <= edge
        u = count * (edge - x) / L
    u = numpy.where(inside, u, 0.0)
    i = numpy.floor(u).astype(numpy.int64)  # int() of a non-negative value
    frac = u - i
    saturated = i >= count
    i = numpy.minimum(i, count - 1)
    interp = profile[i] * (1.0 - frac) + profile[i + 1] * frac
    return numpy.where(inside, numpy.where(saturated, profile[count], interp), 0.0)


def sample_coordinates(n: int, dx: float, shift: int, wraps: bool = True) -> numpy.ndarray:
    """Coordinates of one Yee sub-lattice along an axis of n cells centred on 0: the doubled
    integer -n + 2 i + shift times dx/2, with index 0 of the integer lattice (shift 0) taking
    the seam copy's index n on a wrapping axis."""
    index = numpy.arange(n, dtype=numpy.int64)
    if shift == 0 and wraps:
        index[0] = n
    return (-n + 2 * index + shift) * (0.5 * float(dx))


def graded_sigma(grid, layers: Sequence[AbsorbingFace],
                          wraps: Sequence[bool] = (True, True, True)) -> Dict[str, numpy.ndarray]:
    """The six float64 sigma volumes ("Dx".."Bz", C order (nx, ny, nz)) of a set of layers;
    empty for none. Sigma is a sum over the layers of a function of one axis's coordinate, so
    each volume is the broadcast sum of per-axis lines at the component's own sub-lattice.
    ``wraps`` is False on a metallic axis."""
    if not layers:
        return {}
    resolution = 1.0 / float(grid.dx)
    lines = {}
    for layer in layers:
        if grid.is_invariant(layer.axis):
            raise ValueError(f"axis {layer.axis} is invariant: it takes no absorber")
        n = grid.shape[layer.axis]
        half_length = 0.5 * float(grid.cell_size[layer.axis])
        for shift in (0, 1):
            line = sigma_on_axis(sample_coordinates(n, grid.dx, shift,
                                                          bool(wraps[layer.axis])),
                                    layer, half_length, resolution)
            key = (layer.axis, shift)
            lines[key] = line if key not in lines else lines[key] + line
    shape = tuple(grid.shape)
    out = {}
    for comp in GRADED_COMPONENTS:
        total = numpy.zeros(shape, dtype=numpy.float64)
        for axis in range(3):
            line = lines.get((axis, _SHIFTS[comp][axis]))
            if line is None:
                continue
            b = [1, 1, 1]
            b[axis] = shape[axis]
            total = total + line.reshape(b)
        out[comp] = total
    return out
