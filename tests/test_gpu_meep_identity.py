"""MEEP identity on the GPU: each lifted capsule (the reference's Tier A keep set), pasted into
the HIP engine through the C ABI and stepped there, holds MEEP 1.35.0's stored words of all
twelve components at every compared step (steps 0-20 one at a time, then every tenth of the
run through replayed graphs).

The cases: ``a`` and ``b``, 2-D 200x200 with a 20-cell UPML inside metallic walls (MEEP's
default boundary) and a Gaussian Ez point source, ``b`` with per-component inverse epsilon;
``ex.material-dispersion``, one periodic cell with two Lorentz poles; ``ex.absorber-1d``,
400 cells along z between metallic walls, an mp.Absorber (conductivity on D and B) and five
Drude/Lorentz poles, which runs the general instantiation.
"""

import pytest

from meep_gpu import capsule
from tests import capsule_cases as cc

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("case", cc.CASES)
def test_the_engine_steps_meep_s_words(case):
    state, digests, currents = cc.load(case)
    print(f"meep identity {case}: {state.shape}, {len(state.poles)} poles, "
          f"pml {state.pml is not None}, metallic {state.metallic}", flush=True)
    run = capsule.LiftedRun(state, currents)
    try:
        got = cc.replay_digests(run, digests)
    finally:
        run.close()
    assert set(got) == set(digests.steps)
    assert {step: diff for step, diff in got.items() if diff} == {}


def test_the_replay_entry_point():
    """``capsule.replay`` on a fixture folder: every compared step clean."""
    print("capsule.replay ex.material-dispersion", flush=True)
    out = capsule.replay(cc.case_dir("ex.material-dispersion"))
    assert out and all(not diff for diff in out.values())
