"""Polarization (integrated) sources on the GPU where their replaced words are read beyond
their own cells: an odd Y mirror over periodic boundaries with sources on the images of its
near and far ghost rows, a metallic x wall with a box straddling Ey's wall row, and an
off-diagonal chi1inv row (whose neighbour reads take the replaced words). Every stored word
against the oracle (which replaces D in place and runs MEEP's boundary passes again), consult
by consult and replayed from captured graphs; and the folded run against the unfolded one."""

import numpy
import pytest

from tests import test_folded_polarization as fp

pytestmark = pytest.mark.gpu


def _run(drv, steps: int):
    if steps < 128:
        for _ in range(steps):
            drv.step()
    else:
        drv.run(steps=steps)
    drv.synchronize()


def _driver(case: str, steps: int):
    if case == "mirror":
        return fp.mirror_driver(True, steps)
    if case == "wall":
        return fp.wall_driver(steps)
    # real storage with an off-diagonal Ex-Ey row, a pole and two polarization sources
    from tests import driver_bloch as db

    return db.driver("xy_products_offdiag_integrated", complex_fields=False, k=(0.0, 0.0, 0.0))


@pytest.mark.parametrize("steps", [30, 300])
@pytest.mark.parametrize("case", ["mirror", "wall", "offdiag"])
def test_polarization_sources_match_the_oracle(case, steps):
    print(f"polarization sources, {case} x{steps}", flush=True)
    drv = _driver(case, steps)
    _run(drv, steps)
    ref = fp.oracle_of(drv, steps)
    for _ in range(steps):
        ref.step()
    for key, want in ref.stored_volumes().items():
        got = drv.sim.download(key)
        assert got.view(numpy.uint32).tobytes() == want.view(numpy.uint32).tobytes(), (case, key)
    if case == "wall":
        assert numpy.all(drv.sim.download("Ey")[0] == 0.0)  # the wall row stays cleared
    drv.close()


def test_folded_polarization_run_equals_the_unfolded_run():
    """The stored half of the folded run against the unfolded run with the sources' images
    (rows 1-6 of the fold are rows 6-11 of the 12-row axis), within float32 rounding."""
    steps = 300
    print(f"polarization sources, folded against unfolded x{steps}", flush=True)
    out = []
    for folded in (True, False):
        drv = fp.mirror_driver(folded, steps)
        _run(drv, steps)
        out.append({c: drv.sim.download(c) for c in ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz")})
        drv.close()
    folded, full = out
    scale = max(float(numpy.max(numpy.abs(v))) for v in full.values())
    assert scale > 0
    for c in folded:
        assert numpy.max(numpy.abs(folded[c][:, 1:7, :] - full[c][:, 6:12, :])) <= 1e-5 * scale, c
