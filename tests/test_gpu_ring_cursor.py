"""Reading a monitor mid-run and stepping on (GPU).

A flux read runs the float64 transform over the samples still buffered and empties the ring;
the samples after it must still be transformed in the order they were taken. Here a run is
read after 150 steps (22 samples buffered), after 50 more (fewer than one ring) and after 300
more (two ring periods replayed from a captured graph, then a partial ring), alone and as a
batch member, and every read equals one read of the same total run to float64 rounding and
the float64 oracle to 1e-9 of the spectrum's peak.
"""

import numpy
import pytest

from tests import workload_runs as wr

pytestmark = pytest.mark.gpu

SCHEDULE = (150, 50, 300)


def _spectra(drv):
    return [m.get_flux_spectrum() for m in drv.flux_monitors]


def _close(a, b, rel):
    for x, y in zip(a, b):
        peak = numpy.max(numpy.abs(y))
        assert peak > 0
        assert numpy.max(numpy.abs(x - y)) <= rel * peak


def test_read_step_read_equals_one_read_and_the_oracle():
    from meep_gpu.driver import FdtdDriver

    wl = wr.small("sphere")
    total = sum(SCHEDULE)
    drv = FdtdDriver.from_workload(wl)
    drv.freeze(planned_steps=total + 1)
    done = 0
    for n in SCHEDULE:
        print(f"ring cursor: step {n} more, then read at step {done + n}", flush=True)
        drv.run(steps=n)
        done += n
        got = _spectra(drv)
        single = wr.hip_run(wl, done)
        _close(got, _spectra(single), 1e-12)
        single.close()
        _, planes = wr.oracle_run(wl, done, reference_precision=False)
        _close(got, [p.flux(wl.grid.dx, "c128") for p in planes], 1e-9)
    drv.close()


def test_batch_flush_then_step_on_equals_one_read():
    from meep_gpu.batch import BatchedRun
    from meep_gpu.driver import FdtdDriver

    wls = [wr.small("sphere"), wr.small("sphere_box")]
    drivers = []
    for wl in wls:
        d = FdtdDriver.from_workload(wl)
        d.freeze(planned_steps=401)
        drivers.append(d)
    batch = BatchedRun(drivers)
    print("ring cursor: batch of two, 150 steps, flush, 250 more", flush=True)
    batch.run(150)
    batch.flush()
    batch.run(250)
    batch.synchronize()
    batch.close()
    for wl, d in zip(wls, drivers):
        alone = wr.hip_run(wl, 400)
        _close(_spectra(d), _spectra(alone), 1e-12)
        alone.close()
        d.close()


def test_member_of_a_live_batch_is_not_freed():
    """Closing a member while its batch is alive is refused on the host, and the engine itself
    defers the release of a member destroyed through the C ABI until the batch ends: the batch
    keeps stepping valid memory."""
    from meep_gpu.batch import BatchedRun
    from meep_gpu.driver import FdtdDriver

    wl = wr.small("sphere")
    a, b = FdtdDriver.from_workload(wl), FdtdDriver.from_workload(wl)
    for d in (a, b):
        d.freeze(planned_steps=101)
    batch = BatchedRun([a, b])
    with pytest.raises(RuntimeError):
        a.close()
    b.sim.lib.fdtd_sim_destroy(b.sim.handle)  # deferred: the batch still holds it
    b.sim.handle = None
    print("deferred release: the batch steps on", flush=True)
    batch.run(100)
    batch.synchronize()
    batch.close()  # releases b's device state
    alone = wr.hip_run(wl, 100)
    for comp in ("Ez", "Hy"):
        assert numpy.array_equal(a.sim.download(comp), alone.sim.download(comp))
    alone.close()
    a.close()
