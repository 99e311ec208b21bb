"""Mirror folds on the GPU: every stored word against the oracle, eager, replayed from captured
graphs and stepped inside a batch beside unfolded members.

The recorded reference digest of ``folded_3d`` (an even x and an odd y mirror over periodic
boundaries) is reproduced in test_gpu_parity with the other fixture cases. The variants here
have no recorded digest: ``folded_metallic_3d`` folds two axes over metallic boundaries (an odd
full count on z), ``folded_conductive_pole_3d`` an odd periodic x fold and a metallic z fold
with D and B conductivities and a Lorentz pole.
"""

import types

import numpy
import pytest

from oracle import cases

pytestmark = pytest.mark.gpu

FOLDED = ("folded_3d", "folded_metallic_3d", "folded_conductive_pole_3d")


def _same_words(got, ref, label):
    for key, value in ref.stored_volumes().items():
        assert got[key].view(numpy.uint32).tobytes() == value.view(numpy.uint32).tobytes(), \
            f"{label}: {key}"


@pytest.mark.parametrize("name", FOLDED)
def test_folded_case_matches_the_oracle_over_forty_steps(name):
    from tests import hip_cases

    print(f"case {name} x40", flush=True)
    state = cases.build_state(name)
    sim = hip_cases.hip_simulation(state)
    ref = cases.oracle_sim(cases.build_state(name))
    for _ in range(40):
        sim.step_h()
        sim.step_d()
        ref.step()
    _same_words(hip_cases.hip_volumes(sim, state), ref, name)
    sim.close()


@pytest.mark.parametrize("name", ("folded_3d", "folded_metallic_3d"))
def test_folded_graph_replay_matches_the_oracle(name):
    """Two captured 128-step sequences (the fold kernels inside them) and a partial chunk."""
    from tests import hip_cases

    steps = 300
    print(f"case {name} x{steps} replayed", flush=True)
    state = cases.build_state(name)
    sim = hip_cases.hip_simulation(state)
    sim.run(steps)
    sim.synchronize()
    ref = cases.oracle_sim(cases.build_state(name))
    for _ in range(steps):
        ref.step()
    _same_words(hip_cases.hip_volumes(sim, state), ref, name)
    sim.close()


def test_folded_members_step_in_a_batch_beside_unfolded_ones():
    from meep_gpu.batch import BatchedRun
    from tests import hip_cases

    steps = 150
    names = ("folded_3d", "pml_3d", "folded_conductive_pole_3d", "folded_metallic_3d")
    members, states = [], []
    for name in names:
        state = cases.build_state(name)
        sim = hip_cases.hip_simulation(state)
        members.append(types.SimpleNamespace(sim=sim, grid=sim.grid))
        states.append(state)
    print(f"batch {names} x{steps}", flush=True)
    batch = BatchedRun(members)
    batch.run(steps)
    batch.synchronize()
    batch.close()
    for name, state, member in zip(names, states, members):
        ref = cases.oracle_sim(cases.build_state(name))
        for _ in range(steps):
            ref.step()
        _same_words(hip_cases.hip_volumes(member.sim, state), ref, f"{name} in the batch")
        member.sim.close()
