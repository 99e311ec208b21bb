"""Polarization (integrated) sources on mirror-folded and walled grids: the pieces the engine
subtracts (``meep_gpu.sources.polarization_boxes``: each box off its ghost and wall rows, plus
its images on the ghost rows) against the reference's ``f_minus_p``, which subtracts the box and
runs the D boundary passes again (reference sources.py:1697-1747, restated by the oracle).

Random D volumes (some words exact zeros, so signed zeros are compared too), random boxes that
touch ghost, image and wall rows; the two are compared word for word at every cell."""

import types

import numpy
import pytest

from meep_gpu.sources import DepositBox, polarization_boxes
from oracle import engine as oe

SHAPE = (7, 9, 8)
CASES = {
    # an odd Y mirror over periodic boundaries (near and far ghost rows)
    "y_periodic_odd": dict(folds=(None, (-1, 2, 5), None), metal=(False, False, False)),
    # an even X mirror over metallic boundaries beside a metallic z wall
    "x_metallic_even_z_wall": dict(folds=((1, 1, None), None, None), metal=(False, False, True)),
    # two folded axes (corner ghosts image through both planes) and a y wall
    "xz_folds_y_wall": dict(folds=((-1, 2, 4), None, (1, 1, None)), metal=(False, True, False)),
}


def _boxes(rng, n):
    out = []
    for comp in ("Ex", "Ey", "Ez"):
        for _ in range(n):
            lo = tuple(int(rng.integers(0, 3)) if rng.uniform() < 0.6
                       else int(rng.integers(SHAPE[a] - 4, SHAPE[a] - 1)) for a in range(3))
            cnt = tuple(int(rng.integers(1, min(4, SHAPE[a] - lo[a]) + 1)) for a in range(3))
            re = rng.normal(size=cnt)
            im = rng.normal(size=cnt)
            re[rng.uniform(size=cnt) < 0.2] = 0.0
            out.append(DepositBox(comp, lo, cnt, re, im, integrated=True))
    return out


def _deposit(box: DepositBox) -> oe.PointDeposit:
    ii, jj, kk = numpy.meshgrid(*[numpy.arange(l, l + c) for l, c in zip(box.lo, box.cnt)],
                                indexing="ij")
    flat = numpy.ravel_multi_index((ii.reshape(-1), jj.reshape(-1), kk.reshape(-1)), SHAPE)
    return oe.PointDeposit(box.component, flat.astype(numpy.int64),
                           numpy.asarray(box.amp_re, dtype=numpy.float64).reshape(-1),
                           numpy.asarray(box.amp_im, dtype=numpy.float64).reshape(-1),
                           lambda t: 0.37 - 1.21j, True)


def _state(spec, rng):
    sim = oe.OracleSim(SHAPE, 0.05, 0.1)
    sim.set_metallic(spec["metal"])
    sim.set_folding(spec["folds"])
    for c in "xyz":
        d = rng.normal(size=SHAPE).astype(numpy.float32)
        d[rng.uniform(size=SHAPE) < 0.1] = 0.0
        sim.v["D" + c][...] = d
    # D as update_E finds it: the sub-step's boundary passes have run
    sim.fill_folds("D", far=False)
    sim.zero_metal("D")
    sim.fill_folds("D", far=True)
    return sim


def _engine_words(spec, seed, boxes, image_sign=1.0):
    """Each piece's dipole subtracted at its own cells only (the engine's update_E read)."""
    sim = _state(spec, numpy.random.default_rng(seed))
    grid = types.SimpleNamespace(shape=SHAPE, folds=spec["folds"])
    for box in boxes:
        for piece in polarization_boxes(grid, box, spec["metal"]):
            sim.add_deposit(_deposit(piece))
    if image_sign != 1.0:  # a planted mistake: every image piece with the wrong sign
        for dep in sim.deposits:
            dep.amp_re = dep.amp_re * numpy.where(_is_image(dep, spec), image_sign, 1.0)
    sim.deposit(False, 0.0, integrated=True)
    return {c: sim.v["D" + c].copy() for c in "xyz"}


def _is_image(dep, spec):
    """Whether a piece's cells lie on a ghost row (an image piece)."""
    c = "xyz".index(dep.component[1])
    idx = numpy.unravel_index(dep.index, SHAPE)
    ghost = numpy.zeros(dep.index.shape, dtype=bool)
    for a, fold in enumerate(spec["folds"]):
        if fold is None:
            continue
        if a != c:
            ghost |= idx[a] == 0
        elif fold[1] == 2:
            ghost |= idx[a] == SHAPE[a] - 1
    return ghost


def _reference_words(spec, seed, boxes):
    """MEEP's f_minus_p: every box subtracted, then the boundary passes again."""
    sim = _state(spec, numpy.random.default_rng(seed))
    for box in boxes:
        sim.add_deposit(_deposit(box))
    sim.f_minus_p_begin(False, 0.0)
    return {c: sim.v["D" + c].copy() for c in "xyz"}


def _same(a, b):
    return all(a[c].view(numpy.uint32).tobytes() == b[c].view(numpy.uint32).tobytes()
               for c in "xyz")


@pytest.mark.parametrize("name", sorted(CASES))
@pytest.mark.parametrize("seed", [1, 2, 3])
def test_the_engine_s_pieces_give_meep_s_f_minus_p(name, seed):
    spec = CASES[name]
    boxes = _boxes(numpy.random.default_rng(100 + seed), 4)
    want = _reference_words(spec, seed, boxes)
    got = _engine_words(spec, seed, boxes)
    for c in "xyz":
        assert got[c].view(numpy.uint32).tobytes() == want[c].view(numpy.uint32).tobytes(), \
            (name, seed, c)


@pytest.mark.parametrize("name", sorted(CASES))
def test_planted_mistakes_move_the_words(name):
    """The images' parity flipped, or the box kept on its ghost and wall rows: both differ."""
    spec = CASES[name]
    boxes = _boxes(numpy.random.default_rng(7), 6)
    want = _reference_words(spec, 5, boxes)
    assert _same(_engine_words(spec, 5, boxes), want)
    assert not _same(_engine_words(spec, 5, boxes, image_sign=-1.0), want)
    # the boxes as given, no clipping and no images
    sim = _state(spec, numpy.random.default_rng(5))
    for box in boxes:
        sim.add_deposit(_deposit(box))
    sim.deposit(False, 0.0, integrated=True)
    assert not _same({c: sim.v["D" + c] for c in "xyz"}, want)


def test_an_unfolded_open_grid_keeps_the_box_whole():
    grid = types.SimpleNamespace(shape=SHAPE, folds=(None, None, None))
    box = DepositBox("Ez", (0, 0, 0), (2, 3, 1), numpy.ones((2, 3, 1)), numpy.zeros((2, 3, 1)),
                     integrated=True)
    pieces = polarization_boxes(grid, box, (False, False, False))
    assert len(pieces) == 1 and pieces[0].lo == box.lo and pieces[0].cnt == box.cnt


def mirror_driver(folded: bool, steps: int):
    """An odd Y mirror over periodic boundaries with two polarization sources touching the
    near and far ghost rows' images (folded), or the whole cell with each source's mirror
    image added (unfolded): Ez, even about a plane not normal to z, times the declared -1, has
    a negated image; Ey, odd about its own plane, times -1, keeps its sign."""
    from meep_gpu.driver import FdtdDriver
    from meep_gpu.sources import GaussianSource

    drv = FdtdDriver((1.6, 2.4, 1.6), 5.0, symmetry=(("Y", -1),) if folded else (),
                     max_steps=steps + 8)
    drv.setup_pml({"z": 2})
    fn = GaussianSource(1.2, 0.5)
    for comp, (x, y, z), image_sign in (("Ez", (0.13, 0.2, 0.21), -1.0),
                                        ("Ey", (-0.31, 1.1, -0.17), 1.0)):
        drv.add_volume_source(comp, (x, y, z), (0.0, 0.0, 0.0), fn, amplitude=0.6 - 0.8j,
                              is_integrated=True)
        if not folded:
            drv.add_volume_source(comp, (x, -y, z), (0.0, 0.0, 0.0), fn,
                                  amplitude=image_sign * (0.6 - 0.8j), is_integrated=True)
    return drv


def wall_driver(steps: int):
    """A metallic x wall with a polarization source straddling the wall row of Ey (its integer
    lattice along x puts row 0 on the wall, which MEEP clears after the dipole is subtracted)."""
    from meep_gpu.driver import FdtdDriver
    from meep_gpu.sources import GaussianSource

    drv = FdtdDriver((1.6, 1.2, 1.6), 5.0, boundaries={"x": "metallic"}, max_steps=steps + 8)
    drv.setup_pml({"z": 2})
    amps = numpy.array([[[0.7]], [[-0.4]]])  # rows 0 (the wall) and 1 along x
    drv.add_source(DepositBox("Ey", (0, 2, 3), (2, 1, 1), amps, 0.5 * amps, integrated=True),
                   GaussianSource(1.2, 0.5))
    drv.add_volume_source("Ez", (0.3, -0.1, 0.1), (0.0, 0.0, 0.0), GaussianSource(1.0, 0.6))
    return drv


def oracle_of(drv, steps: int):
    """The oracle's real run of a real-storage driver (part 0 of its two-part build)."""
    from tests import driver_bloch as db

    return db.oracle(drv, steps).re


def _mirror_run(folded: bool, steps: int):
    drv = mirror_driver(folded, steps)
    ref = oracle_of(drv, steps)
    for _ in range(steps):
        ref.step()
    return drv.grid, ref.stored_volumes()


def test_the_folded_oracle_run_is_the_unfolded_run():
    """The mirror's defining invariant on the oracle: each owned stored row of the folded run
    equals the unfolded run's row at the same place (stored row j of either Yee shift sits at
    doubled coordinate 2j - 2 + shift above the plane, full row r at 2r + shift - 12 for the
    12-cell axis: r = j + 5), polarization sources on the ghost rows' images included."""
    steps = 60
    fgrid, folded = _mirror_run(True, steps)
    ugrid, full = _mirror_run(False, steps)
    assert fgrid.shape[1] == 8 and ugrid.shape[1] == 12
    comps = ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz")
    scale = max(float(numpy.max(numpy.abs(full[c]))) for c in comps)
    assert scale > 0
    for c in comps:
        got = folded[c][:, 1:7, :]
        want = full[c][:, 6:12, :]
        assert numpy.max(numpy.abs(got - want)) <= 1e-6 * scale, c
    assert numpy.max(numpy.abs(folded["Ez"][:, 1:7, :])) > 1e-3 * scale
