"""Host-side guard: a simulation stepped by a live batch cannot be closed (CPU, no GPU calls)."""

import weakref

import pytest


class _Batch:
    handle = object()


def test_closing_a_member_of_a_live_batch_is_refused():
    from meep_gpu.engine import HipSimulation

    batch = _Batch()
    sim = HipSimulation.__new__(HipSimulation)
    sim.handle = None
    sim._batch = weakref.ref(batch)
    assert sim.batch() is batch
    with pytest.raises(RuntimeError, match="live BatchedRun"):
        sim.close()
    batch.handle = None  # the batch closed: its members can be
    assert sim.batch() is None
    sim.close()


def test_a_collected_batch_does_not_hold_its_members():
    from meep_gpu.engine import HipSimulation

    batch = _Batch()
    sim = HipSimulation.__new__(HipSimulation)
    sim.handle = None
    sim._batch = weakref.ref(batch)
    del batch
    assert sim.batch() is None
    sim.close()
