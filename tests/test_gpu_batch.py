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


@pytest.mark.parametrize("rows", [2, 4, 8])
def test_the_d_pair_s_halo_tiles_reproduce_the_recorded_digests(rows, golden_digests):
    """The D pair's box rows over LDS-staged halo tiles (the row below and the column left of
    each 256 / rows x rows tile staged once) step the same words as runs of cells: the
    recorded digests of the lossless cases with an absorber (3-D, and 2-D with a pole)."""
    from meep_gpu.autotune import install
    from meep_gpu.engine import get_launch_pair, get_launch_tile
    from oracle import cases
    from tests import hip_cases

    saved = (get_launch_pair(0), get_launch_pair(1))
    install(((1, 0, 17, True), (1, 0, 17, True)), tile=rows)
    try:
        assert get_launch_tile() == rows
        for name in ("pml_3d", "dispersive_pml_2d", "pml_2d"):
            print(f"halo tiles of {rows} rows, case {name}", flush=True)
            digest, _ = hip_cases.hip_case_digest(cases.build_state(name))
            assert digest == golden_digests["case_sha256"][name], (rows, name)
    finally:
        install(saved, tile=0)


def test_halo_tiles_in_a_batch_step_each_member_as_runs_do():
    """The synthetic members (edges 28-44, half with a Lorentz pole) stepped by one batch
    whose D pair runs 32 x 8 halo tiles, replayed from graphs, against each alone over runs."""
    import dataclasses

    from meep_gpu.batch import BatchedRun
    from meep_gpu.driver import FdtdDriver

    steps = 300
    alone, together = [], []
    for wl in _members():
        a = FdtdDriver.from_workload(wl)
        a.freeze(planned_steps=steps + 1)
        a.run(steps=steps)
        a.synchronize()
        alone.append(a)
        b = FdtdDriver.from_workload(wl)
        b.freeze(planned_steps=steps + 1)
        together.append(b)
    print(f"batch with halo tiles x{steps}", flush=True)
    batch = BatchedRun(together)
    cfg = batch.config()
    batch.set_config(dataclasses.replace(cfg, d=(1, 0, 17, True, 1, 8)))
    assert batch.config().d[5] == 8
    batch.run(steps)
    batch.synchronize()
    batch.close()
    for a, b in zip(alone, together):
        _same_words(a, b)
    for d in alone + together:
        d.close()


def _members():
    from meep_gpu.batch import synthetic_batch

    return synthetic_batch(count=5, seed=7, edge_range=(28, 44), steps=300, nfreq=6,
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
    """Captured 128-step sequences (re-captured when a call starts at another ring phase; one
    run's first sequence starts at step 0, where update_H has not yet run) against one eager
    launch per step."""
    from meep_gpu.driver import FdtdDriver
    from meep_gpu.engine import set_graphs
    from tests.workload_runs import small

    steps = 600
    runs = []
    for graphs, chunks in ((False, (steps,)), (True, (70, 300, 230)), (True, (256, 344))):
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
    _same_words(runs[0], runs[1])
    _same_words(runs[0], runs[2])
    for d in runs:
        d.close()


def test_timed_steps_are_the_steps_run_takes():
    """fdtd_run_timed (eager chunks with events around every pair) against fdtd_run (a
    captured sequence, a replay, eager steps): the same words, and sane per-launch means."""
    from meep_gpu.driver import FdtdDriver
    from tests.workload_runs import small

    steps = 300
    runs = []
    for timed in (False, True):
        drv = FdtdDriver.from_workload(small("sphere_box"))
        sim = drv.freeze(planned_steps=steps + 1)
        if timed:
            sim.run(20)
            total, h_ms, d_ms = sim.run_timed(steps - 20)  # chunks of 128, 128 and 24
            assert 0 < h_ms and 0 < d_ms and (h_ms + d_ms) * (steps - 20) <= total * 1.001
        else:
            sim.run(steps)
        sim.synchronize()
        runs.append(drv)
    _same_words(*runs)
    for d in runs:
        d.close()


def test_each_member_steps_the_same_alone_and_in_the_batch():
    from meep_gpu.batch import BatchedRun
    from meep_gpu.driver import FdtdDriver
    from meep_gpu.engine import set_graphs

    steps = 300
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
    batch.run(20)
    batch.run(steps - 20)  # a captured 128-step sequence, replayed, then eager steps
    batch.synchronize()
    batch.close()
    for a, b in zip(alone, together):
        assert b.step_count == steps
        _same_words(a, b)
    for d in alone + together:
        d.close()


def _inv_driver(kind, steps):
    """A small PML run whose inverse epsilon is stored as float words (kind 'float': more than
    65536 distinct), 16-bit indices ('u16') or byte indices ('u8')."""
    from meep_gpu import workloads
    from meep_gpu.driver import FdtdDriver

    # 44^3 = 85,184 cells: enough for more than 65536 distinct words
    wl = workloads.sphere(edge=44, resolution=4.0, pml_cells=6, radius_cells=8.0, nfreq=4,
                          planes="one", steps=steps, source_x=-2.0, plane_x=2.0)
    drv = FdtdDriver.from_workload(wl)
    rng = numpy.random.default_rng({"float": 1, "u16": 2, "u8": 3}[kind])
    if kind == "float":
        inv = {c: rng.uniform(0.3, 1.0, wl.grid.shape).astype(numpy.float32) for c in wl.inv}
    elif kind == "u16":
        table = rng.uniform(0.3, 1.0, 3000).astype(numpy.float32)
        inv = {c: table[rng.integers(0, 3000, wl.grid.shape)] for c in wl.inv}
    else:
        inv = wl.inv
    drv.set_inverse_epsilon(inv)
    drv.freeze(planned_steps=steps + 1)
    return drv


def test_every_inverse_epsilon_storage_steps_the_same_alone_and_mixed_in_a_batch():
    from meep_gpu.batch import BatchedRun

    steps = 90
    kinds = ("float", "u16", "u8")
    alone = []
    for k in kinds:
        d = _inv_driver(k, steps)
        d.run(steps=steps)
        d.synchronize()
        alone.append(d)
    assert [d.sim.inverse_epsilon_bytes()[0] for d in alone] == [4, 2, 1]
    mixed = [_inv_driver(k, steps) for k in kinds]
    batch = BatchedRun(mixed)
    batch.run(steps)
    batch.synchronize()
    batch.close()
    for a, b in zip(alone, mixed):
        _same_words(a, b)
    for d in alone + mixed:
        d.close()
