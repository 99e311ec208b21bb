"""The HIP engine, through the driver API, against MEEP's recorded output for the reference's
two 2-D examples, and against the oracle word for word (GPU). Tolerances as the oracle's test:
2e-5 on the final Ez norms, 1e-4 of the spectrum peak on every flux value."""

import numpy
import pytest

from tests.meep_examples import build, centred_ez_norms, recorded, spectrum_error

pytestmark = pytest.mark.gpu


def run_engine(name):
    from meep_gpu.driver import FdtdDriver

    ex = build(name)
    spec = ex.spec
    drv = FdtdDriver(cell_size=tuple(spec["cell"]), resolution=spec["resolution"], dimensions=2)
    drv.set_inverse_epsilon(ex.inv)
    for term, sigma in ex.poles:
        drv.add_susceptibility(term, sigma)
    drv.setup_pml(int(round(spec["pml"] * spec["resolution"])))
    drv.add_source(ex.boxes, ex.source)
    fl = spec["flux"]
    mons = [drv.add_flux_region(r["center"], r["size"], fl["fcen"], fl["df"], fl["nfreq"],
                                decimation=spec["decimation"]) for r in fl["regions"]]
    drv.freeze(planned_steps=spec["steps"] + 1)
    drv.run(steps=spec["steps"])
    drv.synchronize()
    return drv, mons


@pytest.mark.parametrize("name", ["slab_flux_2d", "dispersive_slab_2d"])
def test_engine_reproduces_meeps_recorded_fields_and_spectra(name):
    want = recorded()[name]["meep"]
    print(f"example {name}", flush=True)
    drv, mons = run_engine(name)
    try:
        mx, l2 = centred_ez_norms(drv.get_array("Ez"))
        assert abs(mx / want["ez_max_abs"] - 1) < 2e-5
        assert abs(l2 / want["ez_l2_norm"] - 1) < 2e-5
        for mon, ref in zip(mons, want["flux"]):
            assert spectrum_error(mon.get_flux_spectrum(), ref) < 1e-4
    finally:
        drv.close()


@pytest.mark.parametrize("name", ["slab_flux_2d", "dispersive_slab_2d"])
def test_engine_fields_equal_the_oracles_and_spectra_agree_to_float64(name):
    from tests.test_oracle_meep_examples import run_oracle

    print(f"example {name}", flush=True)
    drv, mons = run_engine(name)
    try:
        (omx, ol2), oracle_spectra = run_oracle(name)
        assert centred_ez_norms(drv.get_array("Ez")) == (omx, ol2)
        for mon, ref in zip(mons, oracle_spectra):
            got = mon.get_flux_spectrum()
            assert numpy.allclose(got, ref, rtol=1e-9, atol=1e-12 * numpy.abs(ref).max())
    finally:
        drv.close()
