"""The one-rank chunk division's per-chunk conductivity (CPU): the product's restatement
(meep_gpu/chunks.py) and the oracle's (oracle/chunks.py) pick the same cells, and the block case
the reference fixture records keeps its conductivity in exactly the chunks it reaches."""

import numpy
import pytest

from meep_gpu.chunks import IYEE, ChunkDivision
from meep_gpu.grid import Grid
from oracle import chunks as ochunks


def _block(shape, lo, hi, seed):
    rng = numpy.random.default_rng(seed)
    sigma = numpy.zeros(shape, dtype=numpy.float32)
    sl = tuple(slice(a, b) for a, b in zip(lo, hi))
    sigma[sl] = rng.uniform(0.1, 2.0, size=sigma[sl].shape).astype(numpy.float32)
    return sigma


@pytest.mark.parametrize("cell, dims, faces, metallic, lo, hi", [
    ((2.0, 2.0, 0.0), 2, ((2, 2), (2, 2), (0, 0)), (True, True, False), (8, 8, 0), (16, 16, 1)),
    ((2.0, 1.5, 1.25), 3, ((2, 3), (0, 2), (2.5, 2)), (False, True, False), (3, 0, 4), (9, 5, 10)),
    ((1.5, 1.5, 1.5), 3, ((2, 2), (2, 2), (2, 2)), (False, False, False), (0, 0, 0), (12, 12, 12)),
    ((1.5, 1.25, 0.0), 2, ((0, 0), (2, 2), (0, 0)), (True, False, False), (0, 6, 0), (2, 10, 1)),
])
def test_product_and_oracle_pick_the_same_cells(cell, dims, faces, metallic, lo, hi):
    grid = Grid(resolution=8, cell_size=cell, dimensions=dims)
    division = ChunkDivision.one_rank(grid, faces, tuple(not m for m in metallic))
    sigma = _block(grid.shape, lo, hi, 7)
    for name in ("Dx", "Dy", "Dz", "Bx", "By", "Bz"):
        mine = division.conductive_mask(sigma, IYEE[name])
        theirs = ochunks.conductive_cells(grid.shape, grid.invariant, metallic, faces, sigma,
                                          IYEE[name])
        assert numpy.array_equal(mine, theirs), name


def test_the_block_case_keeps_its_conductivity_in_four_chunks():
    """conductive_chunks_pml_2d: the block in the upper quadrant reaches the interior chunk,
    the x-high and y-high edge chunks and their corner (array_path_oracle.py:440-446)."""
    grid = Grid(resolution=8, cell_size=(2.0, 2.0, 0.0), dimensions=2)
    faces = ((2, 2), (2, 2), (0, 0))
    division = ChunkDivision.one_rank(grid, faces, (False, False, True))
    assert division.count == 9
    sigma = _block(grid.shape, (8, 8, 0), (16, 16, 1), 3)
    owner = division.owner_volume(IYEE["Dz"])
    mask = division.conductive_mask(sigma, IYEE["Dz"])
    kept = sorted({int(k) for k in numpy.unique(owner[mask])})
    assert len(kept) == 4
    assert not mask[:4, :4].any() and mask[-1, -1].all()


def test_product_and_oracle_agree_on_random_divisions():
    """Random cells, layer thicknesses (overlapping ones included), walls and sparse sigmas:
    every component lattice keeps its conductivity in the same cells under both rules."""
    rng = numpy.random.default_rng(20261204)
    for trial in range(40):
        dims = 2 if trial % 3 == 0 else 3
        n = [int(rng.integers(4, 15)) for _ in range(3)]
        cell = (n[0] / 8, n[1] / 8, n[2] / 8 if dims == 3 else 0.0)
        grid = Grid(resolution=8, cell_size=cell, dimensions=dims)
        faces = tuple((float(rng.choice([0, 1, 2, 2.5, 3, 6])), float(rng.choice([0, 1, 2, 3, 6])))
                      if not grid.is_invariant(a) else (0.0, 0.0) for a in range(3))
        metallic = tuple(bool(rng.random() < 0.4) and not grid.is_invariant(a) for a in range(3))
        division = ChunkDivision.one_rank(grid, faces, tuple(not m for m in metallic))
        sigma = numpy.where(rng.random(grid.shape) < 0.03,
                            rng.uniform(0.1, 2.0, grid.shape), 0.0).astype(numpy.float32)
        for name in ("Dx", "Dy", "Dz", "Bx", "By", "Bz"):
            mine = division.conductive_mask(sigma, IYEE[name])
            theirs = ochunks.conductive_cells(grid.shape, grid.invariant, metallic, faces, sigma,
                                              IYEE[name])
            assert numpy.array_equal(mine, theirs), (trial, name, faces, metallic)
