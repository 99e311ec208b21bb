"""The driver on complex storage on the GPU: ``FdtdDriver(force_complex_fields=True,
k_point=..., symmetry=..., boundaries=...)`` stepped consult by consult (``step``) and by
captured graphs (``run``), every stored word of both parts against the oracle's two parts;
and a real run equal to the real part of the same run on complex storage."""

import numpy
import pytest

from tests import driver_bloch as db

pytestmark = pytest.mark.gpu


def _same_parts(drv, ref, label):
    for part, osim in enumerate(ref.parts):
        for key, want in osim.stored_volumes().items():
            got = drv.sim.download(key, part)
            assert got.view(numpy.uint32).tobytes() == want.view(numpy.uint32).tobytes(), \
                (label, part, key)


@pytest.mark.parametrize("steps", [24, 300])
@pytest.mark.parametrize("name", sorted(db.RUNS))
def test_driver_complex_run_matches_the_oracle(name, steps):
    print(f"driver complex {name} x{steps}", flush=True)
    drv = db.driver(name)
    if steps < 128:
        for _ in range(steps):
            drv.step()
    else:
        drv.run(steps=steps)
    drv.synchronize()
    ref = db.oracle(drv, steps)
    for _ in range(steps):
        ref.step()
    _same_parts(drv, ref, name)
    ez = drv.get_array("Ez")
    assert ez.dtype == numpy.complex64
    assert numpy.array_equal(ez.real, ref.re.v["Ez"]) and numpy.array_equal(ez.imag, ref.im.v["Ez"])
    assert numpy.any(ez.imag != 0)
    # MEEP's get_array of the same words: the Bloch window's phases, a fold's images
    from meep_gpu.chunks import ChunkDivision
    from meep_gpu.readback import meep_array

    def oracle_read(name):
        out = numpy.empty(drv.grid.shape, dtype=numpy.complex64)
        out.real, out.imag = ref.re.stored_volumes()[name], ref.im.stored_volumes()[name]
        return out

    division = ChunkDivision.one_rank(drv.grid, drv.sim.pml.faces, drv.wraps)
    for comp in ("Ez", "Hx", "By"):
        got = drv.get_meep_array(comp)
        want = meep_array(oracle_read, comp, drv.grid, drv.wraps, division, True, steps,
                          phases=drv.bloch_phases)
        assert got.dtype == numpy.complex64 and got.shape == want.shape
        assert got.view(numpy.uint32).tobytes() == want.view(numpy.uint32).tobytes(), comp
        if drv.grid.folds[1] is not None:  # the whole cell along the fold, images unfolded
            assert got.shape[1] == drv.grid.shape_full[1]
    drv.close()


def test_a_real_run_is_the_real_part_of_the_complex_run():
    steps = 150
    print(f"driver real vs complex x{steps}", flush=True)
    cplx = db.driver("xy_products", k=(0.0, 0.0, 0.0), pml=2)
    real = db.driver("xy_products", complex_fields=False, k=(0.0, 0.0, 0.0), pml=2)
    cplx.run(steps=steps)
    real.run(steps=steps)
    for name in ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz", "Dx", "By"):
        got = cplx.get_array(name)
        want = real.get_array(name)
        assert got.real.view(numpy.uint32).tobytes() == want.view(numpy.uint32).tobytes(), name
    assert numpy.any(cplx.get_array("Ez").imag != 0)
    cplx.close()
    real.close()


@pytest.mark.parametrize("steps", [40, 300])
def test_flux_of_the_complex_field_matches_the_oracle(steps):
    """A flux plane on complex storage: each part's transform on its own part, read as the
    complex field's transform (real part's + i times imaginary part's); the spectrum holds the
    two parts' cross terms. 1e-9 of the peak: both transforms sum in float64, in other orders."""
    print(f"driver complex flux x{steps}", flush=True)
    drv = db.driver("xy_products", flux=True)
    drv.run(steps=steps)
    drv.synchronize()
    got = drv.flux_monitors[0].get_flux_spectrum()
    ref = db.oracle(drv, steps)
    planes = db.oracle_planes(drv)
    db.step_oracle(ref, planes, steps)
    _same_parts(drv, ref, "flux run")
    want = db.complex_flux(planes[0], drv.grid.dx)
    peak = numpy.max(numpy.abs(want))
    assert peak > 0
    assert numpy.max(numpy.abs(got - want)) <= 1e-9 * peak
    # the cross terms matter: the two parts' own fluxes do not add up to it
    own = sum(p.flux(drv.grid.dx) for p in planes[0])
    assert numpy.max(numpy.abs(own - want)) > 1e-3 * peak
    drv.close()
