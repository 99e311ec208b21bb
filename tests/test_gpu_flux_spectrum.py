"""The flux spectrum reduced on the device against the host's sum over the downloaded
transforms (GPU): the same float64 terms summed in another order, so they agree to 1e-13 of
the spectrum's peak; with MEEP's ladder weights (a region placed off the lattice) and without."""

import numpy
import pytest

from tests import workload_runs as wr

pytestmark = pytest.mark.gpu


def _agree(a, b):
    peak = numpy.max(numpy.abs(b))
    assert peak > 0
    assert numpy.max(numpy.abs(a - b)) <= 1e-13 * peak


def test_device_spectrum_equals_the_host_sum_on_the_box_planes():
    drv = wr.hip_run(wr.small("sphere_box"), 150)
    for mon in drv.flux_monitors:
        _agree(mon.get_flux_spectrum(), mon.get_flux_spectrum_host())
    drv.close()


def test_device_spectrum_with_ladder_weights():
    from meep_gpu.driver import FdtdDriver

    wl = wr.small("sphere")
    drv = FdtdDriver.from_workload(wl)
    # off the lattice on every axis: two weighted sample planes and a weighted edge row
    mon = drv.add_flux_region((1.13, 0.07, -0.21), (0.0, 3.3, 2.9), frequencies=[0.1, 0.15, 0.2])
    assert mon.cell_weights is not None and len(mon.normal_planes) == 2
    drv.freeze(planned_steps=161)
    drv.run(steps=160)
    drv.synchronize()
    _agree(mon.get_flux_spectrum(), mon.get_flux_spectrum_host())
    drv.close()
