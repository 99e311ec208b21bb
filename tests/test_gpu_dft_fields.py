"""DFT field monitors on the GPU against the oracle: the same run, each component's float64
transform over its region to 1e-9 of its peak (the sums differ only in association), for
cell-centred and Yee-site regions of E and H components in 3-D boxes and a flat region, with
an absorber and with a Lorentz pole; and the default region, the absorber's interior."""

import numpy
import pytest

from tests import workload_runs as wr

pytestmark = pytest.mark.gpu


def _monitors(drv, wl):
    freqs = wl.planes[0].frequencies
    L = wl.grid.shape[0] * wl.grid.dx  # a cube
    at = lambda *v: tuple(x * L for x in v)  # noqa: E731
    return [
        # five components: two engine regions
        drv.add_dft_monitor(freqs, ("Ex", "Hy", "Ez", "Hz", "Ey"), at(0.04, -0.03, 0.02),
                            at(0.33, 0.28, 0.22)),
        # Yee sites, reaching into the absorber
        drv.add_dft_monitor(freqs[:3], ("Ez", "Hx", "Hz"), at(-0.22, 0.11, 0.0),
                            at(0.44, 0.1, 0.9), yee_grid=True),
        # flat along z between two centre planes
        drv.add_dft_monitor(freqs, ("Ex", "Hy"), at(0.033, 0.022, 0.0413), at(0.22, 0.17, 0.0)),
    ]


@pytest.mark.parametrize("name", ["sphere", "slab"])
def test_dft_field_regions_match_the_oracle(name):
    from meep_gpu.driver import FdtdDriver

    wl = wr.small(name)
    steps = 150
    drv = FdtdDriver.from_workload(wl)
    mons = _monitors(drv, wl)
    print(f"{name}: regions {[(m.lo, m.cnt, m.components, m.yee_grid) for m in mons]}",
          flush=True)
    drv.freeze(planned_steps=steps + 1)
    drv.run(steps=steps)
    drv.synchronize()
    sim, planes = wr.oracle_run(wl, steps, reference_precision=False, dft_monitors=mons)
    regions = planes[len(wl.planes):]
    for mon, region in zip(mons, regions):
        assert region.components == mon.components
        for q, comp in enumerate(mon.components):
            want = region.acc64[q].reshape((mon.num_frequencies,) + mon.cnt)
            got = numpy.stack([mon.get_dft(comp, f) for f in range(mon.num_frequencies)])
            peak = numpy.max(numpy.abs(want))
            assert peak > 0, (name, comp)
            assert numpy.max(numpy.abs(got - want)) <= 1e-9 * peak, (name, mon.lo, comp)
    # the flat region's reduced array is the weighted sum of its two planes
    flat = mons[2]
    arr = flat.get_dft_array("Hy", 1)
    w0, w1 = flat.empty_axes[2]
    block = flat.get_dft("Hy", 1)
    assert arr.shape == flat.cnt[:2]
    assert numpy.allclose(arr, w0 * block[:, :, 0] + w1 * block[:, :, 1], rtol=1e-14, atol=0)
    drv.close()


def test_the_default_region_is_the_absorbers_interior():
    from meep_gpu.driver import FdtdDriver

    wl = wr.small("sphere")  # 36^3 under a 6-cell absorber
    drv = FdtdDriver.from_workload(wl)
    mon = drv.add_dft_monitor([0.15], ("Ez",))
    drv.freeze(planned_steps=50)
    assert mon.lo == (6, 6, 6) and mon.cnt == (24, 24, 24)
    drv.run(steps=40)
    drv.synchronize()
    assert numpy.max(numpy.abs(mon.get_dft("Ez"))) > 0
    drv.close()
