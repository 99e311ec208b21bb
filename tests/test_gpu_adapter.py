"""The reference side of the seam on the GPU: a live-shaped reference driver (the capsule's
object graph as Python objects, its sources' ``current(t, dt)`` from the recorded table)
planned by ``adapter.hip_plan`` and stepped consult by consult in the reference driver's
order. After each step the words come back into the reference's own arrays, and there they
are MEEP 1.35.0's at every compared step; the source tables are planned short (16 steps) so
that the run extends them on the device."""

import numpy
import pytest

from meep_gpu import adapter, capsule
from tests import capsule_cases as cc

pytestmark = pytest.mark.gpu

CONSULTS = ("step_B", "fill_B", "fill_folded_far_ghosts_B", "update_H", "step_D", "fill_D",
            "fill_folded_far_ghosts_D", "update_E", "update_P")


@pytest.mark.parametrize("case", ["a", "c", "ex.material-dispersion"])
def test_a_reference_driver_steps_through_the_seam_to_meep_s_words(case):
    driver = cc.live_driver(case)
    digests = cc.load(case)[1]
    plan = adapter.hip_plan(driver, planned_steps=16)
    assert plan is not None, adapter.last_refusal()
    fields = driver.fields
    grid = plan.run.grid
    last = max(s for s in digests.steps if s <= 60)
    print(f"seam {case}: {grid.shape}, to step {last}", flush=True)
    try:
        for step in range(1, last + 1):
            for slot in CONSULTS:
                assert plan.dispatch(slot, fields), (step, slot)
            if step not in digests.stored:
                continue
            for comp, want in digests.stored[step].items():
                words = getattr(fields, comp, None)
                if words is None and comp[0] == "H":
                    words = getattr(fields, "B" + comp[1])  # no stored H: H is B
                if words is None:
                    continue  # E formed on demand: the engine's own words are checked below
                words = numpy.asarray(words)
                if any(f is not None for f in grid.folds):
                    words = words[~capsule.ghost_cells(grid, comp)]
                assert capsule.word_digest(words) == want, (case, step, comp)
            assert capsule.compare_step(plan.run, digests, step) == [], (case, step)
        assert plan.planned >= last
    finally:
        plan.close()
