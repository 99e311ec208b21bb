"""MEEP identity on the GPU: each lifted capsule (the reference's Tier A keep and flush sets),
pasted into the HIP engine through the C ABI and stepped there, holds MEEP 1.35.0's stored words
of all twelve components at every compared step (steps 0-20 one at a time, then every tenth of
the run through replayed graphs), and MEEP's sim.get_array words of them (readback).

The cases: ``a`` and ``b``, 2-D 200x200 with a 20-cell UPML inside metallic walls (MEEP's
default boundary) and a Gaussian Ez point source, ``b`` with per-component inverse epsilon;
``ex.material-dispersion``, one periodic cell with two Lorentz poles; ``ex.absorber-1d``,
400 cells along z between metallic walls, an mp.Absorber (conductivity on D and B) and five
Drude/Lorentz poles, which runs the general instantiation; ``c`` and ``c_prime``, 3-D
39x18x19 stored cells of a 39x33x35 cell folded by odd Y and Z mirrors over metallic
boundaries, a PML on both x faces and the high y and z faces, an Hx point source on both
mirror planes; the holey waveguide of MEEP's test_holey_wvg_bands, complex storage with a
Bloch factor of -1 on x and an odd Y fold over periodic boundaries. The get_array read-out
unfolds the mirror images with their parities and carries the Bloch window's phases, in the
C++ runtime MEEP ran with where each case was recorded.
"""

import pytest

from meep_gpu import capsule
from tests import capsule_cases as cc

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("mode", cc.MODES)
@pytest.mark.parametrize("case", cc.CASES)
def test_the_engine_steps_meep_s_words(case, mode):
    """Subnormals kept (libmeep_fdtd.so) and flushed (libmeep_fdtd_flush.so: the float32 and
    float64 denormal modes at flush, against the set MEEP minted under x86 FTZ|DAZ)."""
    state, digests, currents = cc.load(case, mode=mode)
    print(f"meep identity {case} ({mode}): {state.shape}, {len(state.poles)} poles, "
          f"pml {state.pml is not None}, metallic {state.metallic}", flush=True)
    run = capsule.LiftedRun(state, currents, subnormals=mode)
    run.runtime = cc.runtime(case, mode)
    try:
        got, read = {}, {}
        for step in digests.steps:
            run.advance(step - run.step_count)
            got[step] = capsule.compare_step(run, digests, step)
            # MEEP's sim.get_array words of the GPU's words (GpuRunResult.get_array)
            read[step] = capsule.compare_reader(run, digests, step)
    finally:
        run.close()
    assert set(got) == set(digests.steps)
    assert {step: diff for step, diff in got.items() if diff} == {}
    assert {step: diff for step, diff in read.items() if diff} == {}


def test_the_replay_entry_point():
    """``capsule.replay`` on a fixture folder: every compared step clean."""
    print("capsule.replay ex.absorber-1d (flush)", flush=True)
    out = capsule.replay(cc.case_dir("ex.absorber-1d", "flush"))
    assert out and all(not diff for diff in out.values())


def test_meep_s_words_in_a_batch():
    """The four lifted cases stepped together by shared batch launches each keep MEEP's words
    (walls and five poles put the whole batch on the general instantiation; 1-D, 2-D and a
    single cell share its tile list). The batch is read at compared steps by closing it, and
    re-formed to go on."""
    from meep_gpu.batch import BatchedRun

    loaded = {case: cc.load(case) for case in cc.CASES}
    runs = {case: capsule.LiftedRun(state, currents)
            for case, (state, _, currents) in loaded.items()}
    try:
        done = 0
        for target, cases in ((1, cc.CASES), (2, cc.CASES), (7, cc.CASES), (20, cc.CASES),
                              (100, ("a", "b", "ex.material-dispersion")),
                              (200, ("a", "b", "ex.material-dispersion"))):
            print(f"batch of {len(runs)} to step {target}", flush=True)
            batch = BatchedRun(list(runs.values()))
            batch.run(target - done)
            batch.synchronize()
            batch.close()
            done = target
            for case in cases:
                digests = loaded[case][1]
                assert runs[case].step_count == target
                if target in digests.stored:
                    assert capsule.compare_step(runs[case], digests, target) == [], (case, target)
    finally:
        for run in runs.values():
            run.close()
