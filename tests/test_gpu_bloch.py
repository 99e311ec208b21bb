"""Complex storage with Bloch phases on the GPU: the engine's two linked parts against the
oracle's, every stored word of both parts, stepped eagerly in a batch of two and replayed from
captured graphs; phases that are products (not +1 or -1) on x, and on x and y together, with
an off-diagonal Ex-Ey row and a Lorentz pole. The oracle's connection is pinned by the
supercell check (tests/test_bloch.py) and by MEEP's words of the holey-waveguide capsule."""

import math
import types

import numpy
import pytest

from tests import bloch_cases as bc

pytestmark = pytest.mark.gpu

RUNS = {
    "x_quarter_offdiag": dict(seed=3, thetas=(math.pi / 2, None, None), offdiag=True),
    "x_third_pole": dict(seed=4, thetas=(2 * math.pi / 3, None, None), offdiag=False, pole=True),
    "xy_offdiag_pole": dict(seed=5, thetas=(0.7, -1.9, None), offdiag=True, pole=True),
}


@pytest.mark.parametrize("steps", [40, 300])
@pytest.mark.parametrize("name", sorted(RUNS))
def test_linked_parts_step_to_the_oracle_s_words(name, steps):
    from meep_gpu.batch import BatchedRun

    spec = dict(RUNS[name])
    run = bc.make(spec.pop("seed"), **spec)
    print(f"bloch {name} x{steps}", flush=True)
    re, im = bc.hip(run)
    batch = BatchedRun([types.SimpleNamespace(sim=re, grid=re.grid),
                        types.SimpleNamespace(sim=im, grid=im.grid)])
    batch.run(steps)
    batch.synchronize()
    stats = batch.graph_stats()
    batch.close()
    if steps >= 128:
        assert stats[0] >= 1 and stats[1] >= 1
    cplx = bc.oracle(run)
    for _ in range(steps):
        cplx.step()
    for part, (sim, osim) in enumerate(((re, cplx.re), (im, cplx.im))):
        for n in bc.NAMES:
            got = sim.download(n)
            want = osim.v[n]
            assert got.view(numpy.uint32).tobytes() == want.view(numpy.uint32).tobytes(), \
                (name, part, n)
    re.close()
    im.close()


def test_a_linked_part_refuses_to_step_alone():
    from meep_gpu._native import EngineError

    run = bc.make(6)
    re, im = bc.hip(run)
    with pytest.raises(EngineError, match="other part"):
        re.run(1)
    re.close()
    im.close()
