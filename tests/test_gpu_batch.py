"""Batched and launch-shape invariance (GPU): a simulation steps to the same words alone, in a
batch of uneven grids, and under every launch shape the autotuner may pick."""

import numpy
import pytest

pytestmark = pytest.mark.gpu

# (H pair shape, D pair shape): cells per thread, workgroups per CU, XCD remap, NT stores
SHAPES = [((1, 0, True, False),) * 2, ((2, 0, True, True),) * 2, ((4, 0, False, False),) * 2,
          ((2, 8, True, False),) * 2, ((4, 4, True, True),) * 2, ((1, 8, False, True),) * 2,
          ((1, 8, False, True), (2, 0, False, True)), ((4, 0, True, False), (1, 4, False, True)),
          ((1, 8, 18, True), (2, 0, 9, True)), ((2, 4, 5, False),) * 2, ((4, 0, 3, True),) * 2]


def _sid(pair):
    return "_".join(f"v{s[0]}b{s[1]}r{int(s[2])}n{int(s[3])}" for s in pair)  # r: tile order


@pytest.mark.parametrize("shapes", SHAPES, ids=_sid)
def test_every_launch_shape_reproduces_the_recorded_digests(shapes, golden_digests):
    from meep_gpu.autotune import install
    from meep_gpu.engine import get_launch_pair
    from oracle import cases
    from tests import hip_cases

    saved = (get_launch_pair(0), get_launch_pair(1))
    install(shapes)
    try:
        for name in ("pml_3d", "offdiag_pml_3d", "dispersive_pml_2d", "no_pml_3d"):
            print(f"shapes {shapes} case {name}", flush=True)
            digest, _ = hip_cases.hip_case_digest(cases.build_state(name))
            assert digest == golden_digests["case_sha256"][name], (shapes, name)
    finally:
        install(saved)


def _members():
    from meep_gpu.batch import synthetic_batch

    return synthetic_batch(count=5, seed=7, edge_range=(28, 44), steps=150, nfreq=6,
                           lorentz_fraction=0.5)


def _same_words(a, b):
    for comp in ("Ex", "Ey", "Ez", "Dx", "Dy", "Dz", "Bx", "By", "Bz", "Hx", "Hy", "Hz"):
        x, y = a.sim.download(comp), b.sim.download(comp)
        assert x.view(numpy.uint32).tobytes() == y.view(numpy.uint32).tobytes(), comp
    for ma, mb in zip(a.flux_monitors, b.flux_monitors):
        fa, fb = ma.get_flux_spectrum(), mb.get_flux_spectrum()
        assert numpy.array_equal(fa, fb)
        assert numpy.max(numpy.abs(fa)) > 0


def test_graph_replay_steps_the_same_words_as_eager_launches():
    """Captured 64-step sequences (re-captured when a call starts at another ring phase)
    against one eager launch per step."""
    from meep_gpu.driver import FdtdDriver
    from meep_gpu.engine import set_graphs
    from tests.workload_runs import small

    steps = 230
    runs = []
    for graphs, chunks in ((False, (steps,)), (True, (70, 130, 30))):
        set_graphs(graphs)
        try:
            drv = FdtdDriver.from_workload(small("sphere_box"))
            drv.freeze(planned_steps=steps + 1)
            for c in chunks:
                drv.run(steps=c)
            drv.synchronize()
        finally:
            set_graphs(True)
        assert drv.step_count == steps
        runs.append(drv)
    _same_words(*runs)
    for d in runs:
        d.close()


def test_each_member_steps_the_same_alone_and_in_the_batch():
    from meep_gpu.batch import BatchedRun
    from meep_gpu.driver import FdtdDriver
    from meep_gpu.engine import set_graphs

    steps = 150
    alone = []
    set_graphs(False)  # alone: eager launches; the batch below replays graphs
    try:
        for wl in _members():
            drv = FdtdDriver.from_workload(wl)
            drv.freeze(planned_steps=steps + 1)
            drv.run(steps=steps)
            drv.synchronize()
            alone.append(drv)
    finally:
        set_graphs(True)
    together = []
    for wl in _members():
        drv = FdtdDriver.from_workload(wl)
        drv.freeze(planned_steps=steps + 1)
        together.append(drv)
    batch = BatchedRun(together)
    batch.run(100)
    batch.run(steps - 100)
    batch.synchronize()
    batch.close()
    for a, b in zip(alone, together):
        assert b.step_count == steps
        _same_words(a, b)
    for d in alone + together:
        d.close()
