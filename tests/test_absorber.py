"""``mp.Absorber``'s graded conductivity (CPU): MEEP's profile table against the numbers the
reference recorded from MEEP (reference test_absorber.py:44-57: L = 1 at resolution 20 gives
N = 40 and sigma_max = 25.904082296183013; L = 0.7 at resolution 13, N = 18 and
37.005831851690026), the lookup against an independent scalar reading of meepgeom.cpp
:1596-1625, the per-component Yee positions, the seam cell, and the driver's single rounding."""

import math

import numpy
import pytest

from meep_gpu.absorber import (GRADED_COMPONENTS, AbsorbingFace, graded_sigma,
                               profile_table, sample_coordinates, faces_for)
from meep_gpu.grid import Grid


def test_profile_table_is_meeps():
    count, prof = profile_table(1.0, 20.0)
    assert count == 40
    assert prof[-1] == pytest.approx(25.904082296183013, rel=1e-14)
    count, prof = profile_table(0.7, 13.0)
    assert count == 18
    assert prof[-1] == pytest.approx(37.005831851690026, rel=1e-14)
    assert profile_table(2.0, 10.0)[0] == 40  # half pixels, not pixels
    a, b = profile_table(0.70, 13.0), profile_table(0.71, 13.0)
    assert a[0] == b[0] == 18 and a[1][-1] != b[1][-1]  # the thickness is not rounded
    with pytest.raises(ValueError):
        profile_table(0.01, 10.0)


def _scalar_sigma(x, L, half, resolution, side, R=1e-15):
    """meepgeom.cpp:1596-1625 as a scalar loop, read independently of the module."""
    n = int(L / (0.5 / resolution) + 0.5)
    prefac = -math.log(R) / (4.0 * L / 3.0)
    prof = [prefac * (i / n) ** 2 for i in range(n + 1)]
    if side == "high":
        edge = half - L
        if x < edge:
            return 0.0
        u = n * (x - edge) / L
    else:
        edge = L - half
        if x > edge:
            return 0.0
        u = n * (edge - x) / L
    i = int(u)
    if i >= n:
        return prof[n]
    f = u - i
    return prof[i] * (1.0 - f) + prof[i + 1] * f


def test_lookup_matches_a_scalar_reading_at_every_yee_position():
    grid = Grid(resolution=13.0, cell_size=(8.0, 8.0, 0.0), dimensions=2)
    layers = faces_for(0.7, direction=0, resolved=(True, True, False))
    vols = graded_sigma(grid, layers)
    assert set(vols) == set(GRADED_COMPONENTS)
    n = grid.shape[0]
    for comp, shift in (("Dz", 0), ("Dx", 1), ("By", 1), ("Bx", 0)):
        sigma = vols[comp][:, 0, 0]
        for i in range(n):
            index = n if (shift == 0 and i == 0) else i  # the seam copy of index n
            x = (-n + 2 * index + shift) * 0.5 * grid.dx
            want = sum(_scalar_sigma(x, 0.7, 4.0, 13.0, side) for side in ("low", "high"))
            assert sigma[i] == pytest.approx(want, rel=1e-13, abs=1e-15), (comp, i)
        assert numpy.all(vols[comp] == vols[comp][:, :1, :1])  # depends on x alone


def test_the_seam_cell_takes_the_far_faces_sigma():
    """A one-sided high layer: index 0 of the integer lattice is MEEP's copy of index n, at
    +L/2, deep in the layer; the half-integer lattice's index 0 sits at -L/2 + dx/2, outside."""
    grid = Grid(resolution=10.0, cell_size=(6.0, 1.0, 0.0), dimensions=2)
    vols = graded_sigma(grid, [AbsorbingFace(1.0, 0, "high")])
    count, prof = profile_table(1.0, 10.0)
    assert vols["Dz"][0, 0, 0] == prof[count]       # shift 0 along x
    assert vols["Dx"][0, 0, 0] == 0.0               # shift 1 along x
    x = sample_coordinates(grid.shape[0], grid.dx, 0)
    assert x[0] == pytest.approx(3.0) and x[1] == pytest.approx(-2.9)


def test_corners_add_and_the_interior_is_zero():
    grid = Grid(resolution=10.0, cell_size=(4.0, 3.0, 2.0))
    layers = faces_for(0.5)
    assert len(layers) == 6
    vols = graded_sigma(grid, layers)
    for comp, v in vols.items():
        assert numpy.all(v[6:-6, 6:-6, 6:-6] == 0.0), comp
        assert v.dtype == numpy.float64
    x_only = graded_sigma(grid, faces_for(0.5, direction=0))
    y_only = graded_sigma(grid, faces_for(0.5, direction=1))
    assert numpy.array_equal(graded_sigma(grid, faces_for(0.5, direction=0)
                                                   + faces_for(0.5, direction=1))["Dz"],
                             x_only["Dz"] + y_only["Dz"])


def test_the_driver_adds_a_material_sigma_in_double_and_rounds_once():
    from meep_gpu.driver import FdtdDriver

    drv = FdtdDriver((3.0, 2.0, 0.0), 10.0, dimensions=2)
    grid = drv.grid
    material = numpy.full(grid.shape, 0.1, dtype=numpy.float64)
    drv.set_conductivity({"Dz": material})
    drv.set_absorber(0.6, direction=0)
    vols = drv.conductivity_volumes()
    ab = graded_sigma(grid, faces_for(0.6, direction=0, resolved=(True, True, False)))
    assert set(vols) == set(GRADED_COMPONENTS)
    assert numpy.array_equal(vols["Dz"], (material + ab["Dz"]).astype(numpy.float32))
    assert vols["Dz"].dtype == numpy.float32
    with pytest.raises(ValueError):
        drv.set_conductivity({"Ex": material})
