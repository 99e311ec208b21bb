"""Graph replay where flux planes decimate (GPU): periods of the decimations' least common
multiple replayed from a captured graph, the transform passes between replays, give the same
field words as every step launched eagerly, and the same spectra to float64 rounding (the
passes between replays take fewer samples than a full ring, so the float64 sums associate
differently: tolerance 1e-12 of the peak); alone, with MEEP's automatic decimation, and in a
batch of members with different decimations."""

import numpy
import pytest

pytestmark = pytest.mark.gpu


def _driver(decimations, freqs=None):
    from meep_gpu.driver import FdtdDriver
    from tests.workload_runs import small

    drv = FdtdDriver.from_workload(small("sphere"))
    n = drv.grid.shape[0]
    freqs = numpy.linspace(0.1, 0.2, 7) if freqs is None else freqs
    mons = list(drv.flux_monitors)
    for q, dec in enumerate(decimations):
        mons.append(drv.add_flux_monitor(0, (n // 2 + 3 + q, 6, 6), (1, n - 12, n - 12),
                                         frequencies=freqs, decimation=dec))
    return drv, mons


def _same(a, ma, b, mb):
    for comp in ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz", "Dz", "By"):
        x, y = a.get_array(comp), b.get_array(comp)
        assert x.view(numpy.uint32).tobytes() == y.view(numpy.uint32).tobytes(), comp
    for p, q in zip(ma, mb):
        sp, sq = p.get_flux_spectrum(), q.get_flux_spectrum()
        peak = numpy.max(numpy.abs(sp))
        assert peak > 0
        assert numpy.max(numpy.abs(sp - sq)) <= 1e-12 * peak


@pytest.mark.parametrize("decimations", [(3, 5), (0,), (43,)])
def test_decimated_planes_replay_the_eager_words(decimations):
    from meep_gpu.engine import set_graphs

    steps = 700
    print(f"decimations {decimations}", flush=True)
    runs = []
    for graphs in (False, True):
        set_graphs(graphs)
        try:
            drv, mons = _driver(decimations)
            drv.freeze(planned_steps=steps)
            drv.run(steps=steps)
            drv.synchronize()
            runs.append((drv, mons, drv.sim.graph_stats()))
        finally:
            set_graphs(True)
    (a, ma, sa), (b, mb, sb) = runs
    assert sa == (0, 0)
    assert sb[0] >= 1 and sb[1] >= 1, sb  # captured, then replayed
    if decimations == (0,):
        assert mb[-1].decimation == b.automatic_decimation(mb[-1].frequencies) > 1
    _same(a, ma, b, mb)
    a.close()
    b.close()


def test_members_with_different_decimations_replay_in_a_batch():
    from meep_gpu.batch import BatchedRun

    steps = 500
    alone = []
    from meep_gpu.engine import set_graphs

    set_graphs(False)
    try:
        for dec in ((7,), (4, 2)):
            drv, mons = _driver(dec)
            drv.freeze(planned_steps=steps)
            drv.run(steps=steps)
            drv.synchronize()
            alone.append((drv, mons))
    finally:
        set_graphs(True)
    members = []
    for dec in ((7,), (4, 2)):
        drv, mons = _driver(dec)
        drv.freeze(planned_steps=steps)
        members.append((drv, mons))
    batch = BatchedRun([d for d, _ in members])
    batch.run(steps)
    batch.synchronize()
    captures, replays = batch.graph_stats()
    batch.close()
    assert captures >= 1 and replays >= 1
    for (a, ma), (b, mb) in zip(alone, members):
        _same(a, ma, b, mb)
        a.close()
        b.close()
