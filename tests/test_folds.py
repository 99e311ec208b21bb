"""Mirror folds on the CPU: the halved storage, the layer table on a folded axis, and the
oracle's fill passes against the reference's recorded ``folded_3d`` digest with planted
mistakes (each must move the digest)."""

import json
import os

import numpy
import pytest

from oracle import cases, fixture

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _grid(cell, symmetry, metallic=(False, False, False)):
    from meep_gpu.grid import Grid

    return Grid(resolution=8, cell_size=cell, symmetry=symmetry, metallic=metallic)


@pytest.mark.parametrize("full,metallic,stored,far", [
    (12, False, 8, 6), (11, False, 8, 5), (10, False, 7, 5),
    (12, True, 7, None), (11, True, 7, None), (9, True, 6, None)])
def test_halved_storage_and_far_row(full, metallic, stored, far):
    """N - N//2 + 1 cells, one more over periodic boundaries, whose last slot images row
    N - stored + 2 (grid.py:568-575, stepping.py:2361-2390)."""
    g = _grid((full / 8.0, 1.0, 1.0), (("X", -1),), (metallic, False, False))
    assert g.shape[0] == stored and g.shape_full[0] == full
    assert g.folds[0] == (-1, 1 if metallic else 2, far)
    assert g.owned(0) == (stored if metallic else stored - 1)
    # the image of the last slot is the mirror image of its doubled coordinate about the second
    # mirror at doubled N: shift-1 row j sits at 2j - 1 (cell 0 one cell below the plane)
    if far is not None:
        assert 2 * full - (2 * (stored - 1) - 1) == 2 * far - 1


def test_fold_refusals():
    with pytest.raises(ValueError):
        _grid((1.5, 1.0, 1.0), (("X", 2),))
    with pytest.raises(ValueError):
        _grid((1.5, 1.0, 1.0), (("X", 1), ("X", -1)))
    from meep_gpu.grid import Grid

    with pytest.raises(ValueError):
        Grid(resolution=8, cell_size=(1.5, 1.0, 0.0), dimensions=2, symmetry=(("Z", 1),))


def test_a_mirror_plane_carries_no_absorber():
    from meep_gpu.pml import PML

    g = _grid((1.5, 1.25, 1.0), (("X", 1),))
    assert PML(g, 2).faces[0] == (0, 2)  # the scalar means every boundary face
    with pytest.raises(ValueError):
        PML(g, {"x": (1, 2)})


@pytest.mark.parametrize("name", ["folded_3d", "folded_metallic_3d", "folded_conductive_pole_3d"])
def test_product_layer_table_equals_the_oracle_on_folded_axes(name):
    from meep_gpu.grid import Grid
    from meep_gpu.pml import PML

    case = cases.CASE_BY_NAME[name]
    g = Grid(resolution=fixture.RESOLUTION, cell_size=case.cell, courant=fixture.COURANT,
             dimensions=case.dimensions, symmetry=case.symmetry, metallic=case.metallic)
    table = PML(g, case.pml, ln_reflection=fixture.PML_LN_REFLECTION,
                wraps=tuple(not m for m in case.metallic))
    osim = cases.oracle_sim(cases.build_state(name))
    assert g.shape == osim.shape
    for key, vectors in table.vectors.items():
        for ours, theirs in zip(vectors, osim._vec[key][:3]):
            assert ours.tobytes() == theirs.tobytes(), key
        assert numpy.array_equal(table.members[key], osim._vec[key][3].astype(bool)), key


def _recorded(name):
    with open(os.path.join(ROOT, "tests", "golden", "array_path_digests.json")) as fh:
        return json.load(fh)["case_sha256"][name]


def test_oracle_reproduces_the_folded_digest():
    assert cases.run_case("folded_3d") == _recorded("folded_3d")


@pytest.mark.parametrize("mistake", ["even_parity", "near_row_1", "far_row_off",
                                     "no_far_fill", "no_near_mask"])
def test_planted_fold_mistakes_move_the_digest(mistake, monkeypatch):
    from oracle import engine

    real_fill = engine.OracleSim.fill_folds
    real_grid = engine.OracleSim._grid

    def fill(self, field, far):
        if mistake == "no_far_fill" and far:
            return
        folds = self.folds
        if mistake == "even_parity":
            self.folds = tuple(None if f is None else (abs(f[0]), f[1], f[2]) for f in folds)
        if mistake == "far_row_off":
            self.folds = tuple(None if f is None else (f[0], f[1], None if f[2] is None
                                                       else f[2] + 1) for f in folds)
        try:
            if mistake == "near_row_1" and not far:
                for axis, fold in enumerate(self.folds):
                    if fold is None:
                        continue
                    for c, comp in enumerate("xyz"):
                        shift0 = (axis == c) if field == "B" else (axis != c)
                        if shift0:
                            vol = self.v[field + comp]
                            dst = [slice(None)] * 3
                            src = [slice(None)] * 3
                            dst[axis], src[axis] = 0, 1
                            vol[tuple(dst)] = vol[tuple(src)]
                return
            real_fill(self, field, far)
        finally:
            self.folds = folds

    def grid(self):
        g = real_grid(self)
        if mistake == "no_near_mask":  # the ghosts stepped as owned cells: no fold at all
            for a in range(3):
                g.fold[a] = 0
        return g

    monkeypatch.setattr(engine.OracleSim, "fill_folds", fill)
    monkeypatch.setattr(engine.OracleSim, "_grid", grid)
    assert cases.run_case("folded_3d") != _recorded("folded_3d")
