"""The dispatch seam on the GPU: stepping through the driver's consults (one fused launch per
pair, answered by the plan) gives the words ``run`` gives, and the plan reports its launches."""

import numpy
import pytest

pytestmark = pytest.mark.gpu


def test_consult_by_consult_steps_match_run():
    from meep_gpu.driver import FdtdDriver
    from tests.workload_runs import small

    steps = 70
    a = FdtdDriver.from_workload(small("slab"))
    a.freeze(planned_steps=steps + 1)
    for _ in range(steps):
        a.step()
    a.synchronize()
    b = FdtdDriver.from_workload(small("slab"))
    b.freeze(planned_steps=steps + 1)
    b.run(steps=steps)
    b.synchronize()
    assert a.step_count == b.step_count == steps
    for comp in ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz", "Dx", "Bz"):
        x, y = a.get_array(comp), b.get_array(comp)
        assert x.view(numpy.uint32).tobytes() == y.view(numpy.uint32).tobytes(), comp
    for ma, mb in zip(a.flux_monitors, b.flux_monitors):
        assert numpy.array_equal(ma.get_flux_spectrum(), mb.get_flux_spectrum())
    rep = a.fast_path_report()
    assert rep["dispatched"] and rep["launches"] == 2 * steps
    a.close()
    b.close()


def test_a_run_past_the_planned_length_extends_the_source_tables():
    """Planned for 40 steps, then stepped and run to 300 (the tables extended on the device
    twice): the same words as a run planned for 300 from the start, and the source still
    drives after step 40."""
    from meep_gpu.driver import FdtdDriver
    from tests.workload_runs import small

    a = FdtdDriver.from_workload(small("sphere"))
    a.freeze(planned_steps=40)
    for _ in range(45):
        a.step()
    a.run(steps=255)
    a.synchronize()
    b = FdtdDriver.from_workload(small("sphere"))
    b.freeze(planned_steps=300)
    b.run(steps=300)
    b.synchronize()
    assert a.step_count == b.step_count == 300
    for comp in ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz", "Dz", "Bx"):
        x, y = a.get_array(comp), b.get_array(comp)
        assert x.view(numpy.uint32).tobytes() == y.view(numpy.uint32).tobytes(), comp
    assert numpy.max(numpy.abs(b.get_array("Ez"))) > 0
    a.close()
    b.close()


def test_each_simulation_keeps_its_own_launch_configuration():
    """A simulation with a configuration of its own keeps it when the process defaults change,
    one without follows them; both step to the same words (spectra to float64 rounding: the
    vector and matrix-core transforms associate differently)."""
    from meep_gpu.driver import FdtdDriver
    from meep_gpu.engine import LaunchConfig, get_launch_pair, set_launch_pair
    from tests.workload_runs import small

    steps = 300
    a = FdtdDriver.from_workload(small("sphere"))
    b = FdtdDriver.from_workload(small("sphere"))
    a.freeze(planned_steps=steps)
    b.freeze(planned_steps=steps)
    own = LaunchConfig(h=(4, 0, 1, True, 2, 0), d=(1, 4, 0, False, 1, 4), graphs=False,
                       flux_mfma=0)
    a.sim.set_config(own)
    saved = get_launch_pair(0)
    try:
        set_launch_pair(0, 2, 8, 0, False)
        assert a.sim.config() == own
        assert b.sim.config().h[:4] == (2, 8, 0, False)
        a.run(steps=steps)
        b.run(steps=steps)
    finally:
        set_launch_pair(0, *saved)
    a.synchronize()
    b.synchronize()
    assert a.sim.graph_stats() == (0, 0) and b.sim.graph_stats()[0] >= 1
    for comp in ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz", "Dz", "By"):
        x, y = a.get_array(comp), b.get_array(comp)
        assert x.view(numpy.uint32).tobytes() == y.view(numpy.uint32).tobytes(), comp
    for ma, mb in zip(a.flux_monitors, b.flux_monitors):
        sa, sb = ma.get_flux_spectrum(), mb.get_flux_spectrum()
        assert numpy.max(numpy.abs(sa - sb)) <= 1e-12 * numpy.max(numpy.abs(sa))
    a.sim.set_config(None)
    assert a.sim.config() == b.sim.config()
    a.close()
    b.close()
