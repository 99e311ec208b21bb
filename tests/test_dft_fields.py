"""DFT field monitors on the host (CPU): where a MEEP volume registers (the flux ladder, reference
dft.py:604-706), how ``get_dft_array`` reduces flat and invariant axes (reference
from_meep.py:886-955), the grouping of components into engine regions, and the refusals."""

import math

import numpy
import pytest

from meep_gpu.dft import DftFieldsMonitor, axis_cells, boundary_weight_ladder
from meep_gpu.grid import Grid


class _Regions:
    """add_dft_region / flux_read of a device simulation, on host arrays."""

    def __init__(self):
        self.regions = []

    def add_dft_region(self, lo, cnt, comps, omegas, scale, dec, yee):
        rng = numpy.random.default_rng(len(self.regions) + 7)
        shape = (len(comps), len(omegas)) + tuple(cnt)
        self.regions.append((tuple(comps), rng.standard_normal(shape)
                             + 1j * rng.standard_normal(shape)))
        return len(self.regions) - 1

    def flux_read(self, rid):
        return self.regions[rid][1].copy()


def test_a_flat_axis_registers_the_bracketing_centre_planes_with_their_weights():
    grid = Grid(resolution=10.0, cell_size=(4.0, 3.0, 2.0))
    mon = DftFieldsMonitor.from_region(grid, ("Ex", "Hz"), (0.3, -0.2, 0.137),
                                       (1.0, 0.6, 0.0), [0.5, 0.6])
    for a, (c, sz) in enumerate(zip((0.3, -0.2, 0.137), (1.0, 0.6, 0.0))):
        first, w = axis_cells(grid.shape[a], grid.dx, c - sz / 2, c + sz / 2)
        if sz == 0:
            assert (mon.lo[a], mon.cnt[a]) == (first, 2)
            assert numpy.allclose(mon.empty_axes[a], w) and math.isclose(sum(w), 1.0)
            # 20 cells along z: centre planes at z = 0.05 and 0.15, and 0.137 lies 0.87 of a
            # cell above the first
            assert numpy.allclose(mon.empty_axes[a], (0.13, 0.87))
        else:
            assert (mon.lo[a], mon.cnt[a]) == (first, w.size)
            assert a not in mon.empty_axes
    # on a centre plane the far plane's weight is zero and only one plane registers
    on = DftFieldsMonitor.from_region(grid, ("Ey",), (0.0, 0.0, 0.05), (0.4, 0.4, 0.0), [0.5])
    assert on.cnt[2] == 1 and on.empty_axes[2] == (1.0,)


def test_an_invariant_axis_is_one_cell_and_must_be_flat():
    grid = Grid(resolution=10.0, cell_size=(4.0, 3.0, 0.0), dimensions=2)
    mon = DftFieldsMonitor.from_region(grid, ("Ez",), (0.0, 0.0, 0.0), (1.0, 0.0, 0.0), [0.5])
    assert mon.cnt[2] == 1 and mon.invariant == (False, False, True)
    with pytest.raises(ValueError):
        DftFieldsMonitor.from_region(grid, ("Ez",), (0.0, 0.0, 0.0), (1.0, 0.0, 0.5), [0.5])


def test_get_dft_array_sums_flat_axes_under_their_weights_and_drops_invariant_ones():
    grid = Grid(resolution=10.0, cell_size=(4.0, 3.0, 0.0), dimensions=2)
    mon = DftFieldsMonitor.from_region(grid, ("Ez", "Hx"), (0.3, 0.137, 0.0), (1.0, 0.0, 0.0),
                                       [0.4, 0.5, 0.6])
    mon.attach(_Regions())
    for comp in ("Ez", "Hx"):
        block = mon.get_dft(comp, 2)
        assert block.shape == mon.cnt
        w0, w1 = mon.empty_axes[1]
        want = w0 * block[:, 0, 0] + w1 * block[:, 1, 0]
        got = mon.get_dft_array(comp, 2)
        assert got.shape == (mon.cnt[0],)
        assert numpy.allclose(got, want, rtol=1e-14, atol=0)


def test_components_go_four_to_an_engine_region():
    mon = DftFieldsMonitor(("Ex", "Hy", "Ez", "Hz", "Ey"), (1, 2, 3), (4, 5, 6), [0.1, 0.2],
                           0.05)
    sim = _Regions()
    mon.attach(sim)
    assert [r[0] for r in sim.regions] == [("Ex", "Hy", "Ez", "Hz"), ("Ey",)]
    for comp in mon.components:
        rid, slot = mon._slots[comp]
        assert numpy.array_equal(mon.get_dft(comp, 1), sim.regions[rid][1][slot, 1])
    assert math.isclose(mon.scale, 0.05 / math.sqrt(2 * math.pi))


def test_refusals():
    with pytest.raises(ValueError):
        DftFieldsMonitor(("Ex", "Ex"), (0, 0, 0), (1, 1, 1), [0.1], 0.05)
    with pytest.raises(ValueError):
        DftFieldsMonitor(("Dx",), (0, 0, 0), (1, 1, 1), [0.1], 0.05)
    with pytest.raises(ValueError):
        DftFieldsMonitor(("Ex",), (0, 0, 0), (1, 1, 1), [0.1], 0.05, decimation=0)
    grid = Grid(resolution=10.0, cell_size=(4.0, 3.0, 2.0))
    with pytest.raises(ValueError):  # a Yee-lattice interpolation is not placed
        DftFieldsMonitor.from_region(grid, ("Ex",), (0, 0, 0.137), (1, 1, 0), [0.5],
                                     yee_grid=True)
    mon = DftFieldsMonitor(("Ex",), (0, 0, 0), (1, 1, 1), [0.1], 0.05)
    with pytest.raises(RuntimeError):
        mon.get_dft("Ex")
    mon.attach(_Regions())
    with pytest.raises(ValueError):
        mon.get_dft("Ey")
    with pytest.raises(IndexError):
        mon.get_dft("Ex", 1)


def test_the_ladder_for_a_flat_axis_is_a_linear_interpolation():
    # the closed form behind the weights: at a point f of the way between two centre planes
    for frac in (0.1, 0.37, 0.5, 0.93):
        x = (3 + 0.5 + frac) / 10.0  # between centres 3.5 and 4.5 (cells of 0.1)
        is_d, ie_d, (s0, s1, e0, e1) = boundary_weight_ladder(x, x, 10.0)
        assert (is_d, ie_d) == (7, 9)
        assert math.isclose(s0, 1 - frac) and math.isclose(e0, frac)
