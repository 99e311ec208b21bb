"""Complex storage with a Bloch phase on the oracle: a cell whose x axis carries
exp(i*theta) steps as ``copies`` cells of the same run with plain periodic wraps and
Bloch-periodic data (theta = 2*pi/copies), to float32 rounding (the supercell's data are
rounded once from exp(i*m*theta)*f, the cell's connections round at every step); planted
mistakes in the connection (conjugate for the word above, the parts' signs swapped, the phase
dropped) move it by O(1). With an off-diagonal Ex-Ey row (its pair and product connected) and
with a Lorentz pole."""

import math

import numpy
import pytest

from tests import bloch_cases as bc

STEPS = 24


def _gap(run, copies, mutate=None):
    cplx = bc.oracle(run)
    if mutate:
        mutate(cplx)
    a, b = bc.supercell(run, copies)
    for _ in range(STEPS):
        cplx.step()
    bc.step_pair(a, b, STEPS)
    worst = 0.0
    for name in bc.NAMES:
        got = cplx.re.v[name].astype(numpy.complex128) + 1j * cplx.im.v[name].astype(
            numpy.complex128)
        want = bc.bloch_window(a, b, name, run)
        scale = max(numpy.abs(want).max(), 1e-30)
        worst = max(worst, float(numpy.abs(got - want).max() / scale))
    return worst


@pytest.mark.parametrize("copies,offdiag,pole", [(4, True, False), (3, False, True),
                                                 (4, True, True)])
def test_a_bloch_cell_steps_as_its_periodic_supercell(copies, offdiag, pole):
    run = bc.make(11 + copies, thetas=(2 * math.pi / copies, None, None), offdiag=offdiag,
                  pole=pole)
    assert _gap(run, copies) < 2e-5


@pytest.mark.parametrize("mistake", ["conj_everywhere", "part_signs", "no_phase"])
def test_planted_connection_mistakes_are_caught(mistake):
    run = bc.make(15, thetas=(2 * math.pi / 4, None, None))

    def mutate(cplx):
        for s in cplx.parts:
            if mistake == "conj_everywhere":
                s.bloch_ph = tuple((a, -b) for a, b in s.bloch_ph)
            elif mistake == "part_signs":
                s.part = 1 - s.part
            else:
                s.bloch_kinds = (1, 0, 0)

    assert _gap(run, 4, mutate) > 1e-2
