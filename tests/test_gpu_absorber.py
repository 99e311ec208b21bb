"""``mp.Absorber`` in place of a PML (GPU): the sphere workload with a graded D and B
conductivity on every face steps to the oracle's words (fields word for word, spectra to 1e-9
of their peak), through the engine's conductive curl forms."""

import numpy
import pytest

from tests import workload_runs as wr

pytestmark = pytest.mark.gpu


def test_absorber_run_matches_the_oracle():
    from meep_gpu.driver import FdtdDriver
    from meep_gpu.dft import FluxMonitor
    from meep_gpu.sources import current_table
    from oracle import workload as ow

    wl = wr.small("sphere")  # 36^3 at resolution 4
    wl.pml_faces = None
    steps = 160
    drv = FdtdDriver.from_workload(wl)
    drv.set_absorber(1.5)  # 6 cells on every face
    sigma = drv.conductivity_volumes()
    assert set(sigma) == {"Dx", "Dy", "Dz", "Bx", "By", "Bz"}
    print("absorber run", flush=True)
    drv.freeze(planned_steps=steps + 1)
    drv.run(steps=steps)
    drv.synchronize()

    grid = wl.grid
    deposits = [(b.component, b.lo, b.cnt, b.amp_re, b.amp_im,
                 current_table(fn, grid.dt, steps + 1, b.magnetic)) for b, fn in wl.sources]
    planes = []
    for p in wl.planes:
        mon = FluxMonitor(p.normal, p.lo, p.cnt, p.frequencies, grid.dx, grid.dt)
        planes.append((p.normal, p.lo, p.cnt, mon.omegas, mon.scale))
    poles = [(t.frequency, t.gamma, t.is_drude, s) for t, s in wl.poles]
    sim, oplanes = ow.build(grid.shape, grid.dt, grid.dx, ((0, 0),) * 3, wl.inv, poles,
                            deposits, planes, grid.invariant)
    sim.set_conductivity(sigma)
    ow.run(sim, oplanes, steps, reference_precision=False)
    for comp in ("Dx", "Dy", "Dz", "Bx", "By", "Bz", "Ex", "Ey", "Ez"):
        got = drv.sim.download(comp)
        want = sim.get_E(comp) if comp[0] == "E" else sim.stored_volumes()[comp]
        assert got.view(numpy.uint32).tobytes() == want.view(numpy.uint32).tobytes(), comp
    for mon, plane in zip(drv.flux_monitors, oplanes):
        got = mon.get_flux_spectrum()
        ref = plane.flux(grid.dx, "c128")
        peak = numpy.max(numpy.abs(ref))
        assert peak > 0 and numpy.max(numpy.abs(got - ref)) <= 1e-9 * peak
    drv.close()


def test_a_source_inside_a_conductor_matches_the_oracle():
    """The Ez point source deposits into a conductive Dz: MEEP scales the current by that
    cell's condinv word in the same double expression (step.cpp:299-307), in the engine's
    general instantiation as in the oracle (fields word for word)."""
    from meep_gpu.driver import FdtdDriver
    from meep_gpu.sources import current_table
    from oracle import workload as ow

    wl = wr.small("sphere")
    wl.planes = []
    steps = 96
    drv = FdtdDriver.from_workload(wl)
    drv.set_conductivity({"Dz": 0.4, "Dx": 0.1})
    sigma = drv.conductivity_volumes()
    grid = wl.grid
    box = wl.sources[0][0]
    assert box.component == "Ez"
    condinv = 1.0 / (1.0 + (0.4 * grid.dt) * 0.5)
    assert numpy.float32(condinv) != numpy.float32(1.0)
    print("source inside a conductor", flush=True)
    drv.freeze(planned_steps=steps + 1)
    drv.run(steps=steps)
    drv.synchronize()

    deposits = [(b.component, b.lo, b.cnt, b.amp_re, b.amp_im,
                 current_table(fn, grid.dt, steps + 1, b.magnetic)) for b, fn in wl.sources]
    poles = [(t.frequency, t.gamma, t.is_drude, s) for t, s in wl.poles]
    from meep_gpu.pml import normalize_faces

    sim, _ = ow.build(grid.shape, grid.dt, grid.dx, normalize_faces(wl.pml_faces, grid), wl.inv,
                      poles, deposits, [], grid.invariant)
    sim.set_conductivity(sigma)
    ow.run(sim, [], steps, reference_precision=False)
    vols = sim.stored_volumes()
    for comp in ("Dx", "Dy", "Dz", "Bx", "By", "Bz", "Ex", "Ey", "Ez"):
        got = drv.sim.download(comp)
        want = sim.get_E(comp) if comp[0] == "E" else vols[comp]
        assert got.view(numpy.uint32).tobytes() == want.view(numpy.uint32).tobytes(), comp
    drv.close()
