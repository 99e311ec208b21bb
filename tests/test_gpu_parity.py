"""HIP engine against the reference's recorded digests and the oracle, word for word (GPU)."""

import numpy
import pytest

from oracle import cases

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", [case.name for case in cases.CASES])
def test_hip_reproduces_recorded_digest(name, golden_digests):
    from tests import hip_cases

    print(f"case {name}", flush=True)
    digest, _ = hip_cases.hip_case_digest(cases.build_state(name))
    assert digest == golden_digests["case_sha256"][name]


@pytest.mark.parametrize("name", ["pml_3d", "dispersive_pml_2d", "offdiag_pml_3d",
                                  "conductive_pml_3d", "offdiag_pml_metallic_3d"])
def test_hip_matches_oracle_over_forty_steps(name):
    """Every stored word after 40 steps; ``offdiag_pml_metallic_3d`` (this build's own case: an
    off-diagonal chi1inv inside metallic walls with a PML on them) has no recorded digest."""
    from tests import hip_cases

    print(f"case {name} x40", flush=True)
    state = cases.build_state(name)
    sim = hip_cases.hip_simulation(state)
    ref = cases.oracle_sim(cases.build_state(name))
    for step in range(40):
        sim.step_h()
        sim.step_d()
        ref.step()
    got = hip_cases.hip_volumes(sim, state)
    for key, value in ref.stored_volumes().items():
        assert got[key].view(numpy.uint32).tobytes() == value.view(numpy.uint32).tobytes(), key
    sim.close()
