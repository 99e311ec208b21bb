"""Flux-data save/load and minus-flux on the GPU, through fdtd_flux_write / fdtd_flux_scale:
a load replaces the transforms word for word (also with samples still buffered), a scale
multiplies them as float64 complex products, and the two-run normalization idiom ends holding
the difference of the runs' transforms (reference dft.py:3261-3316)."""

import numpy
import pytest

pytestmark = pytest.mark.gpu


def _sphere(eps):
    from meep_gpu import workloads

    return workloads.sphere(edge=36, resolution=4.0, pml_cells=6, radius_cells=7.0, eps=eps,
                            fcen=0.15, df=0.1, nfreq=8, planes="one", steps=200,
                            source_x=-2.0)


def _run(wl, steps, **kw):
    from meep_gpu.driver import FdtdDriver
    from meep_gpu.from_meep import run_driver

    drv = FdtdDriver.from_workload(wl)
    return run_driver(drv, steps=steps, **kw)


def _same_data(a, b):
    assert len(a.planes) == len(b.planes)
    for x, y in zip(a.planes, b.planes):
        for side in ("E", "H"):
            for name in getattr(x, side):
                assert numpy.array_equal(getattr(x, side)[name], getattr(y, side)[name]), name


def test_load_replaces_and_scale_multiplies_word_for_word():
    first = _run(_sphere(2.25), 150)
    mon = first.monitors[0]
    data = first.get_flux_data(mon)
    assert numpy.max(numpy.abs(data.planes[0].E["Ey"])) > 0
    second = _run(_sphere(2.25), 50)  # 50 samples still buffered in the ring
    mon2 = second.monitors[0]
    print("load after 50 buffered samples", flush=True)
    mon2.load_dft_data(data)
    _same_data(mon2.get_dft_data(), data)
    z = complex(2.0, -0.5)
    mon2.scale_dfts(z)
    got = mon2.get_dft_data().planes[0]
    for side in ("E", "H"):
        for name, x in getattr(data.planes[0], side).items():
            re = x.real * z.real - x.imag * z.imag
            im = x.real * z.imag + x.imag * z.real
            assert numpy.array_equal(getattr(got, side)[name], re + 1j * im), name
    first.close()
    second.close()


def test_minus_flux_data_subtracts_the_fields():
    from meep_gpu.dft import flux_spectrum

    steps = 200
    empty = _run(_sphere(1.0), steps)
    data = empty.get_flux_data(empty.monitors[0])
    t1 = empty.monitors[0].transforms()
    flux1 = empty.get_flux_spectrum(empty.monitors[0])
    full = _run(_sphere(2.25), steps)
    t2 = full.monitors[0].transforms()
    flux2 = full.get_flux_spectrum(full.monitors[0])
    print("normalization run with the empty run's data loaded minus", flush=True)
    from meep_gpu.driver import FdtdDriver
    from meep_gpu.from_meep import run_driver

    drv = FdtdDriver.from_workload(_sphere(2.25))
    minus = run_driver(drv, steps=steps, minus_flux_data=[(drv.flux_monitors[0], data)])
    t3 = minus.monitors[0].transforms()
    scale = numpy.max(numpy.abs(t2))
    assert numpy.max(numpy.abs(t3 - (t2 - t1))) <= 1e-12 * scale
    mon = minus.monitors[0]
    flux3 = minus.get_flux_spectrum(mon)
    want = flux_spectrum(t2 - t1, mon.area_element, mon.weight, mon.cell_weights)
    assert numpy.allclose(flux3, want, rtol=1e-9, atol=1e-12 * numpy.max(numpy.abs(want)))
    # the fields are subtracted, not the powers
    assert numpy.max(numpy.abs(flux3 - (flux2 - flux1))) > 1e-3 * numpy.max(numpy.abs(flux2))
    for r in (empty, full, minus):
        r.close()
