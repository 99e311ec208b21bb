"""The oracle on the synthetic workloads (CPU): the reference's complex64 flux accumulation
agrees with the float64 one within the 1e-4 flux tolerance, and energy leaves through the
absorber (a decaying field after the source has switched off)."""

import numpy
import pytest

from tests import workload_runs as wr


@pytest.mark.parametrize("name", ["sphere", "slab"])
def test_reference_precision_flux_is_within_tolerance_of_float64(name):
    wl = wr.small(name)
    sim, planes = wr.oracle_run(wl, 120)
    for plane in planes:
        f64 = plane.flux(wl.grid.dx, "c128")
        f32 = plane.flux(wl.grid.dx, "c64")
        peak = numpy.max(numpy.abs(f64))
        assert peak > 0
        assert numpy.max(numpy.abs(f32 - f64)) <= 1e-4 * peak


def test_flux_transform_matches_a_direct_dft_of_the_sampled_fields():
    """The oracle's accumulation equals sum_n exp(i w t_n) f_n dt/sqrt(2 pi) formed directly."""
    import math
    from oracle import workload as ow

    wl = wr.small("sphere")
    sim, planes = wr.oracle_run(wl, 0)
    plane = planes[0]
    samples, times = [], []
    for _ in range(40):
        sim.step()
        plane.update(sim, sim.time)
        samples.append(plane.sample(sim, plane.components[0], sim.get_E(plane.components[0])))
        times.append(sim.time)
    samples = numpy.array(samples, dtype=numpy.float64)
    times = numpy.array(times)
    direct = numpy.einsum("nf,nc->fc", numpy.exp(1j * numpy.outer(times, plane.omegas)),
                          samples) * plane.scale
    assert numpy.allclose(plane.acc64[0], direct, rtol=1e-12, atol=1e-14 * numpy.abs(direct).max())


def test_synthetic_dispersive_member_stays_bounded():
    """The synthetic batch keeps the Lorentz pole out of the UPML: with it inside the layer the
    coupled update grows without bound (x1e4 by step 1500 on this member), without it the
    field after the pulse stays at the pulse's own scale."""
    import numpy

    from meep_gpu.batch import synthetic_batch
    from oracle import workload as ow
    from tests.workload_runs import oracle_run

    members = [w for w in synthetic_batch(count=64, seed=20261203, steps=3000) if w.poles]
    w = min(members, key=lambda m: m.grid.total_cells)
    sim, planes = oracle_run(w, 0, reference_precision=False)
    ow.run(sim, planes, 600, False)
    early = float(numpy.abs(sim.get_E("Ez")).max())
    ow.run(sim, planes, 1400, False)
    late = float(numpy.abs(sim.get_E("Ez")).max())
    assert numpy.isfinite(late) and late < 10 * early
