"""Where a source deposits: the product's placement, the oracle's, the reference's recorded point
counts and the mathematics of MEEP's weights (CPU).

* The point counts the reference records for its fixture sources (``form_classes`` in
  tests/golden/array_path_digests.json: ``<case>/deposit_D/Ez/point`` and the like) are the
  oracle's: a point on a sample row is one sample there, not two with a zero weight.
* The product (``meep_volume_source``) and the oracle place every source the same way: the
  same amplitude words at each cell, in the same order where two lattice images meet, over
  points, sheets and full-width sheets across periodic and metallic axes, in 2-D and 3-D.
* A zero-size direction is a linear interpolation: its weights sum to 1 and their first moment
  is the point's coordinate. An extended direction integrates: its weights sum to the extent in
  cells. A full-width sheet across a periodic axis covers every row with total weight 1 and
  deposits the two images that meet on the seam row separately.
"""

import random

import numpy
import pytest

from meep_gpu.grid import Grid
from meep_gpu.sources import meep_volume_source
from oracle import sources as placement

#: The reference's fixture sources (array_path_oracle.py:447-475): case -> (cell, sources as
#: (component, center, size, integrated, recorded kind))
FIXTURE_SOURCES = {
    "repair_pml_2d": ((1.5, 1.5), [("Ez", (0.0625, 0.0625, 0.0), (0, 0, 0), False, "point"),
                                   ("Hz", (-0.1875, 0.125, 0.0), (0, 0, 0), False, "point")]),
    "sources_pml_2d": ((2.0, 2.0), [("Ez", (-0.8125, 0.0625, 0.0), (0, 0, 0), False, "point"),
                                    ("Ex", (0.0, 0.25, 0.0), (0.75, 0, 0), False, "sheet"),
                                    ("Hz", (0.1875, -0.1875, 0.0), (0, 0, 0), False, "point"),
                                    ("Ey", (0.125, 0.0, 0.0), (0, 0, 0), True, "integrated")]),
    "conductive_source_2d": ((1.5, 1.25), [("Ez", (0.0625, -0.0625, 0.0), (0, 0, 0), False,
                                            "point"),
                                           ("Ex", (0.0, 0.125, 0.0), (0.5, 0, 0), False, "sheet"),
                                           ("Ey", (-0.125, 0.0, 0.0), (0, 0, 0), True,
                                            "integrated")]),
}

_IYEE = placement.IYEE


def _product_points(grid, component, center, size, metallic=(False, False, False)):
    cells, re, im = [], [], []
    ny, nz = grid.shape[1], grid.shape[2]
    for box in meep_volume_source(grid, component, center, size, metallic=metallic):
        for a in range(box.cnt[0]):
            for b in range(box.cnt[1]):
                for c in range(box.cnt[2]):
                    cells.append(((box.lo[0] + a) * ny + box.lo[1] + b) * nz + box.lo[2] + c)
                    re.append(box.amp_re[a, b, c])
                    im.append(box.amp_im[a, b, c])
    return numpy.array(cells, dtype=numpy.int64), numpy.array(re), numpy.array(im)


def _grid2(cell):
    return Grid(resolution=8.0, cell_size=(cell[0], cell[1], 0.0), courant=0.5, dimensions=2)


@pytest.mark.parametrize("case", sorted(FIXTURE_SOURCES))
def test_recorded_point_counts(case, golden_digests):
    cell, sources = FIXTURE_SOURCES[case]
    grid = _grid2(cell)
    for component, center, size, integrated, kind in sources:
        family = "B" if component[0] == "H" else "D"
        key = f"{case}/deposit_{family}/{component}/{kind}"
        recorded = golden_digests["form_classes"][key]
        cells, re, im = placement.deposit_points(grid.shape, grid.invariant, 8.0, component,
                                                 center, size)
        assert cells.size == recorded, key
        assert numpy.all(re != 0) and numpy.all(im == 0), key  # real, nonzero (the mint's)
        mine = _product_points(grid, component, center, size)
        assert numpy.array_equal(mine[0], cells), key
        assert mine[1].tobytes() == re.tobytes(), key


def _random_requests(rng, grid, count):
    out = []
    for _ in range(count):
        component = rng.choice(sorted(_IYEE))
        center, size = [], []
        for d in range(3):
            if grid.is_invariant(d):
                center.append(0.0)
                size.append(0.0)
                continue
            L = grid.shape[d] * grid.dx
            kind = rng.random()
            if kind < 0.35:
                w = 0.0
            elif kind < 0.5:
                w = L  # the whole cell: the periodic seam, or the walls
            else:
                w = rng.uniform(0.0, 0.6 * L)
            c = 0.0 if w == L else rng.uniform(-0.5 * L + w / 2 + grid.dx, 0.5 * L - w / 2 - grid.dx)
            if rng.random() < 0.3:  # onto the lattice (a sample or a midpoint)
                c = round(c / (0.5 * grid.dx)) * 0.5 * grid.dx
            center.append(c)
            size.append(w)
        out.append((component, tuple(center), tuple(size)))
    return out


@pytest.mark.parametrize("dims,metallic", [(2, (False, False, False)), (3, (False, False, False)),
                                           (3, (True, False, True)), (2, (True, True, False))])
def test_product_and_oracle_place_alike(dims, metallic):
    rng = random.Random(20261204 + dims + 2 * sum(metallic))
    cell = (1.5, 1.25, 1.0 if dims == 3 else 0.0)
    grid = Grid(resolution=8.0, cell_size=cell, courant=0.5, dimensions=dims)
    for component, center, size in _random_requests(rng, grid, 60):
        want = placement.deposit_points(grid.shape, grid.invariant, 8.0, component, center, size,
                                        wraps=tuple(not m for m in metallic))
        got = _product_points(grid, component, center, size, metallic)
        label = (component, center, size)
        assert _per_cell(*got) == _per_cell(*want), label


def _per_cell(cells, re, im):
    """Each cell's deposits in the order they are made: what the deposited words depend on
    (points of one pass are distinct cells, so their order among themselves is free)."""
    out = {}
    for c, r, i in zip(cells.tolist(), re.tolist(), im.tolist()):
        out.setdefault(c, []).append((numpy.float64(r).tobytes(), numpy.float64(i).tobytes()))
    return out


def _sample_coordinate(n, dx, iyee, row):
    return (-(n - n % 2) / 2 + row + iyee / 2) * dx


@pytest.mark.parametrize("n", [12, 13])
def test_a_point_interpolates_linearly(n):
    """Along one axis, a zero-size request's weights sum to 1 and place its first moment on the
    coordinate; a request on a sample row is that one row."""
    grid = Grid(resolution=8.0, cell_size=(n / 8.0, 1.0, 0.0), courant=0.5, dimensions=2)
    rng = random.Random(n)
    for trial in range(200):
        component = rng.choice(("Ex", "Ey", "Hz", "Ez"))
        iyee = _IYEE[component][0]
        x = rng.uniform(-0.3 * n / 8.0, 0.3 * n / 8.0)
        if trial % 4 == 0:
            x = _sample_coordinate(n, grid.dx, iyee, rng.randrange(2, n - 2))
        boxes = meep_volume_source(grid, component, (x, 0.1, 0.0), (0.0, 0.0, 0.0))
        rows = {}
        for box in boxes:  # sum out y (its weights sum to 1) and the scale a^2 = 64
            for a in range(box.cnt[0]):
                r = box.lo[0] + a
                rows[r] = rows.get(r, 0.0) + float(box.amp_re[a].sum()) / 64.0
        total = sum(rows.values())
        moment = sum(_sample_coordinate(n, grid.dx, iyee, r) * w for r, w in rows.items())
        assert total == pytest.approx(1.0, abs=1e-12)
        assert moment == pytest.approx(x, abs=1e-12)
        if trial % 4 == 0:
            assert len(rows) == 1


def test_an_extent_integrates_and_a_full_sheet_covers_every_row_once():
    grid = Grid(resolution=8.0, cell_size=(2.0, 1.5, 1.25), courant=0.5)
    rng = random.Random(7)
    for _ in range(100):
        component = rng.choice(sorted(_IYEE))
        w = rng.uniform(0.4, 1.6)
        c = rng.uniform(-1.0 + w / 2, 1.0 - w / 2)
        boxes = meep_volume_source(grid, component, (c, 0.0, 0.0), (w, 1.5, 1.25))
        total = sum(float(b.amp_re.sum()) for b in boxes)
        assert total == pytest.approx(w * 8.0 * (1.5 * 8.0) * (1.25 * 8.0), rel=1e-12)
    # a full-width sheet across periodic y and z: every (y, z) row weight 1 in total, and the
    # rows where two lattice images meet deposit twice
    boxes = meep_volume_source(grid, "Ez", (0.3, 0.0, 0.0), (0.0, 1.5, 1.25))
    cover = numpy.zeros(grid.shape[1:])
    hits = numpy.zeros(grid.shape[1:], dtype=int)
    for b in boxes:
        sl = (slice(b.lo[1], b.lo[1] + b.cnt[1]), slice(b.lo[2], b.lo[2] + b.cnt[2]))
        cover[sl] += b.amp_re.sum(axis=0) / 8.0
        hits[sl] += 1
    assert numpy.allclose(cover, 1.0, atol=1e-12)
    assert hits.max() > 1 and hits.min() == 1
    # the same sheet between metallic y walls: no deposit on the wall rows' far image
    walled = meep_volume_source(grid, "Ez", (0.3, 0.0, 0.0), (0.0, 1.5, 1.25),
                                metallic=(False, True, False))
    rows_y = sorted({b.lo[1] + q for b in walled for q in range(b.cnt[1])})
    assert rows_y == list(range(1, grid.shape[1]))  # row 0 is the low wall, not owned

