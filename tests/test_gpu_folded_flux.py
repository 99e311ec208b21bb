"""Flux and sources on a mirror-folded grid on the GPU, against the unfolded run of the same
symmetric problem (the mirror's defining invariant: the folded run is the unfolded run's
stored half, so its spectra are the unfolded spectra up to float32 rounding; the stated flux
tolerance is 1e-4 of the peak).

An even Y mirror; an Ez Gaussian point source on the mirror plane, and one below it (served
at its mirror image, so the unfolded run carries the pair); an x-normal region straddling the
plane (its image half added), y-normal regions above and below it (the one below is the
image of the one above, its flux negated), all under a PML on the outer faces; DFT field
regions straddling the plane, their arrays over the whole region (the images read back with
each component's parity)."""

import numpy
import pytest

pytestmark = pytest.mark.gpu

CELL = (3.2, 3.2, 2.4)
FREQS = [0.8, 1.0, 1.2]
#: DFT field regions straddling the plane: flat in z (its interpolation weights per centre)
#: and a box
DFT_REGIONS = ((("Ez", "Hx", "Hy"), (0.2, 0.03, 0.0), (1.0, 1.6, 0.0)),
               (("Ey", "Hz"), (0.0, -0.1, 0.1), (0.6, 1.0, 0.6)))
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
    dfts = [drv.add_dft_monitor(FREQS, comps, center=c, size=s) for comps, c, s in DFT_REGIONS]
    drv.run(steps=steps)
    drv.synchronize()
    out = [m.get_flux_spectrum() for m in mons]
    arrays = [m.get_dft_array(comp, 1) for m, (comps, _, _) in zip(dfts, DFT_REGIONS)
              for comp in comps]
    shape = drv.grid.shape
    drv.close()
    return out, shape, arrays


@pytest.mark.parametrize("source_y", [0.0, -0.3])
def test_folded_flux_equals_the_unfolded_run(source_y):
    print(f"folded flux, source at y={source_y}", flush=True)
    folded, fshape, farrays = _run(True, source_y)
    full, ushape, uarrays = _run(False, source_y)
    assert fshape[1] < ushape[1]  # the stored half (with its ghost rows)
    # DFT arrays over the full region, against the largest transform of the run (Hz, the
    # component an Ez dipole leaves at zero, is rounding noise on its own scale)
    scale = max(numpy.max(numpy.abs(u)) for u in uarrays)
    assert scale > 0
    for k, (f, u) in enumerate(zip(farrays, uarrays)):
        assert f.shape == u.shape, (k, f.shape, u.shape)
        assert numpy.max(numpy.abs(f - u)) <= 1e-4 * scale, k
    for k, (f, u) in enumerate(zip(folded, full)):
        peak = numpy.max(numpy.abs(u))
        assert peak > 0
        assert numpy.max(numpy.abs(f - u)) <= 1e-4 * peak, (k, f, u)
    # the region below the plane is the image of the one above: its flux is negated
    assert numpy.max(numpy.abs(folded[2] + folded[1])) <= 1e-4 * numpy.max(numpy.abs(folded[1]))
