"""The dispatch seam's contract on the CPU (reference fastpath.py:4631-4700): slot names, the
identity guard, fused pairs that answer together within one step and not across steps, the
far-fill pass names, and the named refusal of state that is not device-resident."""

import pytest

from meep_gpu.fastpath import (DRIVER_SLOTS, FAR_FILL_PASSES, FUSED_PAIRS, HipStepPlan,
                               last_dispatch_report, plan_fast_path)
from meep_gpu.driver import FdtdDriver


class CountingFields:
    """Stands in for the device-resident simulation: counts the fused launches."""

    def __init__(self):
        self.calls = []

    def step_h(self):
        self.calls.append("H")

    def step_d(self):
        self.calls.append("D")


def test_slots_are_the_reference_drivers():
    assert DRIVER_SLOTS == ("step_B", "fill_B", "update_H", "step_D", "fill_D", "update_E",
                            "update_P")
    assert FAR_FILL_PASSES == {"fill_B": "fill_folded_far_ghosts_B",
                               "fill_D": "fill_folded_far_ghosts_D"}
    covered = [s for pair in FUSED_PAIRS for s in pair]
    assert sorted(covered) == sorted(DRIVER_SLOTS)


def test_one_step_runs_two_launches_and_every_consult_answers_true():
    f = CountingFields()
    plan = HipStepPlan(f)
    for _ in range(3):
        assert all(plan.dispatch(slot, f) for slot in FdtdDriver.CONSULTS)
    assert f.calls == ["H", "D"] * 3
    assert plan.launches[0] == 6


def test_identity_guard_and_unknown_slots_run_nothing():
    f = CountingFields()
    plan = HipStepPlan(f)
    assert plan.dispatch("step_B", CountingFields()) is False
    assert plan.dispatch("zero_metal_B", f) is False
    assert f.calls == []


def test_a_trailing_consult_without_its_lead_in_the_same_step_answers_false():
    f = CountingFields()
    plan = HipStepPlan(f)
    assert plan.dispatch("update_H", f) is False  # the lead never ran
    assert plan.dispatch("step_B", f) is True
    assert plan.dispatch("update_H", f) is True
    # the electric lead is skipped: its trailing consults cannot claim a launch
    assert plan.dispatch("update_E", f) is False
    # a position that does not advance opens a new step and clears the commitment
    assert plan.dispatch("update_H", f) is False
    assert plan.dispatch("fill_folded_far_ghosts_B", f) is False
    assert f.calls == ["H"]


def test_far_fill_passes_answer_with_their_pair():
    f = CountingFields()
    plan = HipStepPlan(f)
    assert plan.dispatch("fill_folded_far_ghosts_D", f) is False
    plan.dispatch("step_B", f)
    plan.dispatch("fill_B", f)
    assert plan.dispatch("fill_folded_far_ghosts_B", f) is True
    plan.dispatch("update_H", f)
    plan.dispatch("step_D", f)
    assert plan.dispatch("fill_folded_far_ghosts_D", f) is True


def test_plan_refuses_host_state_by_name():
    assert plan_fast_path(CountingFields(), None, None) is None
    rep = last_dispatch_report()
    assert rep["dispatched"] is False
    assert any("device-resident" in r for r in rep["reasons"])
