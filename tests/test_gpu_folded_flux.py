"""Flux and sources on a mirror-folded grid on the GPU, against the unfolded run of the same
symmetric problem (the mirror's defining invariant: the folded run is the unfolded run's
stored half, so its spectra are the unfolded spectra up to float32 rounding; the stated flux
tolerance is 1e-4 of the peak).

An even Y mirror; an Ez Gaussian point source on the mirror plane, and one below it (served
at its mirror image, so the unfolded run carries the pair); an x-normal region straddling the
plane (its image half added), y-normal regions above and below it (the one below is the
image of the one above, its flux negated), all under a PML on the outer faces."""

import numpy
import pytest

pytestmark = pytest.mark.gpu

CELL = (3.2, 3.2, 2.4)
FREQS = [0.8, 1.0, 1.2]
REGIONS = (((0.5, 0.0, 0.0), (0.0, 2.0, 1.6)),   # straddles the plane
           ((0.0, 0.6, 0.0), (2.0, 0.0, 1.6)),   # above it
           ((0.0, -0.6, 0.0), (2.0, 0.0, 1.6)))  # below it: the image of the one above


def _run(folded: bool, source_y: float, steps: int = 240):
    from meep_gpu.driver import FdtdDriver
    from meep_gpu.sources import GaussianSource

    drv = FdtdDriver(CELL, 10.0, symmetry=(("Y", 1),) if folded else (), max_steps=steps + 8)
    drv.setup_pml(8)
    fn = GaussianSource(1.0, 0.6)
    drv.add_volume_source("Ez", (-0.4, source_y, 0.1), (0.0, 0.0, 0.0), fn)
    if source_y != 0.0 and not folded:  # the image the mirror implies
        drv.add_volume_source("Ez", (-0.4, -source_y, 0.1), (0.0, 0.0, 0.0), fn)
    mons = [drv.add_flux_region(c, s, frequencies=FREQS) for c, s in REGIONS]
    drv.run(steps=steps)
    drv.synchronize()
    out = [m.get_flux_spectrum() for m in mons]
    shape = drv.grid.shape
    drv.close()
    return out, shape


@pytest.mark.parametrize("source_y", [0.0, -0.3])
def test_folded_flux_equals_the_unfolded_run(source_y):
    print(f"folded flux, source at y={source_y}", flush=True)
    folded, fshape = _run(True, source_y)
    full, ushape = _run(False, source_y)
    assert fshape[1] < ushape[1]  # the stored half (with its ghost rows)
    for k, (f, u) in enumerate(zip(folded, full)):
        peak = numpy.max(numpy.abs(u))
        assert peak > 0
        assert numpy.max(numpy.abs(f - u)) <= 1e-4 * peak, (k, f, u)
    # the region below the plane is the image of the one above: its flux is negated
    assert numpy.max(numpy.abs(folded[2] + folded[1])) <= 1e-4 * numpy.max(numpy.abs(folded[1]))
