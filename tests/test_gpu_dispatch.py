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
