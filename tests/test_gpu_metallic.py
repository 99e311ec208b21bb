"""Metallic (PEC) walls on the GPU against the oracle: 3-D sphere and slab runs with walls on
x and an absorber on y and z, and a sphere with the absorber on x too -- MEEP's PML inside
metallic walls, graded from the walls with no wrap (fields word for word, spectra to 1e-9 of
their peak) -- and the refusal the engine makes by name. Off-diagonal chi1inv inside walls is
checked against the oracle in test_gpu_parity.py."""

import numpy
import pytest

from tests import workload_runs as wr

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name, layer_on_walls", [("sphere", False), ("slab", False),
                                                   ("sphere", True)])
def test_metallic_walls_match_the_oracle(name, layer_on_walls):
    from meep_gpu.dft import FluxMonitor
    from meep_gpu.driver import FdtdDriver
    from meep_gpu.pml import normalize_faces
    from meep_gpu.sources import current_table
    from oracle import workload as ow

    wl = wr.small(name)
    cells = 6
    wl.pml_faces = {"x": cells, "y": cells, "z": cells} if layer_on_walls else \
        {"y": cells, "z": cells}
    steps = 160
    drv = FdtdDriver.from_workload(wl)
    drv.set_boundaries({"x": "metallic"})
    print(f"{name} with metallic x walls, layer on them {layer_on_walls}", flush=True)
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
    sim, oplanes = ow.build(grid.shape, grid.dt, grid.dx, normalize_faces(wl.pml_faces, grid),
                            wl.inv, poles, deposits, planes, grid.invariant)
    sim.set_metallic((True, False, False))
    ow.run(sim, oplanes, steps, reference_precision=False)
    vols = sim.stored_volumes()
    for comp in ("Ex", "Ey", "Ez", "Dx", "Dy", "Dz", "Bx", "By", "Bz", "Hx", "Hy", "Hz"):
        got = drv.sim.download(comp)
        want = vols[comp]
        assert got.view(numpy.uint32).tobytes() == want.view(numpy.uint32).tobytes(), comp
    # the walls hold: tangential D and normal B at x = 0
    assert not numpy.any(drv.sim.download("Dy")[0]) and not numpy.any(drv.sim.download("Bx")[0])
    for mon, plane in zip(drv.flux_monitors, oplanes):
        got = mon.get_flux_spectrum()
        ref = plane.flux(grid.dx, "c128")
        peak = numpy.max(numpy.abs(ref))
        assert peak > 0 and numpy.max(numpy.abs(got - ref)) <= 1e-9 * peak
    drv.close()


def test_what_metallic_walls_refuse():
    """An invariant axis has no wall (the engine's own refusal, beneath the driver's)."""
    from meep_gpu._native import EngineError
    from meep_gpu.engine import HipSimulation
    from meep_gpu.grid import Grid

    sim = HipSimulation(Grid(resolution=4, cell_size=(4.0, 4.0, 0.0), dimensions=2))
    try:
        with pytest.raises(EngineError, match="invariant"):
            sim.set_boundary(2, "metallic")
    finally:
        sim.close()
