"""The reference-side seam adapter on the CPU: a live reference driver read through the
capsule reader gives the capsule's own state, and a duck-typed reference ``Fields`` driven
through ``dispatch(slot, fields)`` gets the reference's contract (fastpath.py:4631-4700):
the identity guard, answers that are never partial, a fused pair's one-step commitment and
the far-fill pass names; the step's words come back after ``update_P``."""

import types

import numpy
import pytest

from meep_gpu import adapter, capsule
from tests import capsule_cases as cc


@pytest.mark.parametrize("case", ["a", "c", "ex.absorber-1d"])
def test_a_live_driver_lifts_to_the_capsule_s_state(case):
    live = adapter.lift_live(cc.live_driver(case))
    read = cc.load(case)[0]
    assert live.shape == read.shape and live.symmetry == read.symmetry
    assert live.metallic == read.metallic and live.step == read.step
    assert set(live.volumes) == set(read.volumes)
    for name, vol in read.volumes.items():
        assert live.volumes[name].tobytes() == vol.tobytes(), name
    if read.pml is not None:
        for key, vectors in read.pml.vectors.items():
            for a, b in zip(vectors, live.pml.vectors[key]):
                assert a.tobytes() == b.tobytes()
            assert numpy.array_equal(read.pml.members[key], live.pml.members[key])
    for a, b in zip(read.sources, live.sources):
        assert numpy.array_equal(a.index, b.index) and numpy.array_equal(a.amps, b.amps)


def test_live_currents_are_asked_at_the_driver_s_times():
    driver = cc.live_driver("c")  # an Hx source: magnetic, at n*dt
    state, digests, recorded = cc.load("c")
    tables = adapter.live_currents(driver, state, 40)
    assert numpy.array_equal(tables[0], recorded[0][:40])
    asked = []
    driver._sources[0].current = lambda t, dt: asked.append(t) or 0j
    driver._sources[0].component = "Ex"
    state.sources[0] = capsule.LiftedSource("Ex", state.sources[0].index, state.sources[0].amps)
    adapter.live_currents(driver, state, 3)
    assert asked == [(n + 0.5) * state.dt for n in range(3)]


def test_what_the_engine_does_not_step_is_refused_before_allocating():
    driver = cc.live_driver("a")
    driver.fields.grid.cylindrical = True
    assert adapter.hip_plan(driver) is None
    assert "cylindrical" in adapter.last_refusal()[0]


class _FakeSim:
    """Counts the fused launches (no GPU): step_h and step_d advance one step between them."""

    def __init__(self):
        self.launches = []
        self.step_count = 0
        self.grid = None

    def step_h(self):
        self.launches.append("H")

    def step_d(self):
        self.launches.append("D")
        self.step_count += 1

    def download(self, name):
        return numpy.full((2, 2, 2), float(self.step_count), dtype=numpy.float32)

    def set_source_currents(self, index, table):
        self.launches.append(("currents", index, len(table)))


def _adapter(sync=True):
    fields = types.SimpleNamespace(Bx=numpy.zeros((2, 2, 2), numpy.float32),
                                   Dz=numpy.zeros((2, 2, 2), numpy.float32), h_separated=False)
    driver = types.SimpleNamespace(fields=fields, _sources=[])
    sim = _FakeSim()
    state = types.SimpleNamespace(step=0, sources=[], dt=0.05)
    run = types.SimpleNamespace(sim=sim, state=state, engine_has=lambda name: True,
                                close=lambda: None)
    return adapter.HipPlanAdapter(driver, run, planned=8, sync=sync), fields, sim


CONSULTS = ("step_B", "fill_B", "fill_folded_far_ghosts_B", "update_H", "step_D", "fill_D",
            "fill_folded_far_ghosts_D", "update_E", "update_P")


def test_the_seam_contract_with_a_duck_typed_fields():
    plan, fields, sim = _adapter()
    other = types.SimpleNamespace()
    # the identity guard: another Fields gets False and nothing runs
    assert not any(plan.dispatch(slot, other) for slot in CONSULTS)
    assert sim.launches == []
    # one step in the driver's order: two fused launches, every consult answered True
    assert all(plan.dispatch(slot, fields) for slot in CONSULTS)
    assert sim.launches == ["H", "D"]
    # the step's words came back after update_P, and H has started
    assert numpy.all(fields.Bx == 1.0) and numpy.all(fields.Dz == 1.0) and fields.h_separated
    # a trailing consult whose lead did not run this step answers False, never half
    plan2, fields2, sim2 = _adapter()
    assert plan2.dispatch("update_H", fields2) is False
    assert plan2.dispatch("fill_folded_far_ghosts_B", fields2) is False
    assert sim2.launches == []
    # a commitment lasts one step: a position that does not advance opens a new step
    assert plan2.dispatch("step_B", fields2) and plan2.dispatch("update_H", fields2)
    assert plan2.dispatch("step_B", fields2)  # a new step: its own H launch
    assert sim2.launches == ["H", "H"]
    assert plan2.dispatch("fill_B", fields2) and plan2.dispatch("update_H", fields2)
    assert plan2.dispatch("update_E", fields2) is False  # step_D has not led in this step


def test_without_sync_nothing_comes_back():
    plan, fields, sim = _adapter(sync=False)
    for slot in CONSULTS:
        plan.dispatch(slot, fields)
    assert numpy.all(fields.Bx == 0.0)
